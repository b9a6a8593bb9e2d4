"""Device-resident H-matrix of the six traction operators.

Build (one GPU, no MPI): cluster tree + block partition on the host
(``cluster.py``), then
  * near-field (non-admissible) blocks: one batched FP64 TDE assembly launch,
    stored in the pool as FP32 (or FP64) -- replaces the per-column ctypes
    loop of ``calc_stressfunc_fromCS/CD`` (src/Hmatrix.py:425-511) run for
    every dense block by ``ACA_worker`` (src/Hmatrix.py:1014-1066);
  * far-field (admissible) blocks: partially pivoted ACA of the stacked
    3m x 2n operator block, one workgroup per block, entries evaluated inside
    the kernel -- replaces ``ACA_plus_full_traceS/D`` (src/Hmatrix.py:581-1011)
    and its row/column evaluators. The reference stops on an absolute
    criterion (eps=0.01, src/Hmatrix.py:1020, :736); this build uses a
    relative Frobenius criterion ||u_k|| ||v_k|| <= eps ||U V||_F.
Apply: ``matvec(v1, v2)`` returns (A1s v1 + A1d v2, A2s v1 + A2d v2,
Bs v1 + Bd v2) in one pass over the stored operators -- the six
``blocks_process_MVM`` calls of one time step (src/QDsim.py:906-917,
src/Hmatrix.py:1445-1538).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .cluster import build_block_partition_device, build_cluster_tree_device
from .mesh import FaultMesh, get_eleVec, receivers
from .tde import hs_records, source_records

DENSE_CHUNK = 32   # columns per dense work item
LR_CHUNK = 16      # rank columns per low-rank work item
UNIT_COLS = 512    # stacked V columns per phase-1 unit
NRHS = (1, 2, 4, 8, 16)
# "f64": every product summed in FP64 (bitwise independent of nrhs); with FP32
# storage and 16 right-hand sides, "mfma" runs every block product on FP32
# matrix cores (FP32 sums per phase-1 unit and per phase-2 wavefront, phase 2
# as a stream of 4-column steps, pq3d_hmat_matvec_mfs) and "mfma_comp" runs
# phase 2 per item group with the items added in compensated FP32 (hi + lo)
ACCUMULATE = ("f64", "mfma", "mfma_comp")
# matrix-core step stream: candidates of the autotuner (HMatrix.tune) and the
# configuration used before tuning
MFS_DEPTHS = (2, 4, 6, 8)     # 4-column steps in flight per wavefront
MFS_WAVES = (1, 2, 4)         # wavefronts per row tile
MFS_ORDERS = ("row", "cost", "rowbal")  # tile launch order: leaf (row) order, XCD-balanced
                                        # by cost, or leaf order in cost-balanced XCD ranges
MFP_DEPTHS = (2, 4)           # phase 1: steps in flight per wavefront
MFP_WAVES = (2, 4)            # phase 1: wavefronts per unit
MFS_DEFAULT = {"depth": 4, "waves": 1, "order": "row", "p1_depth": 4, "p1_waves": 4}
VP_UNIT_COLS = 2048           # stacked V columns per phase-1 unit of the V panels
VP_ALIGN = 32                 # panel bases in floats (128 B: every full 64-rank chunk starts a cache line)
STEP_CNT_SHIFT = 28           # step record word 3: (columns - 1) << 28 | weight row


def _sync(dev):
    if dev.type == "cuda":
        torch.cuda.synchronize(dev)


@dataclass
class HStats:
    n: int
    n_dense: int
    n_lowrank: int
    dense_entries: int
    lowrank_entries: int
    ranks_mean: float
    ranks_max: int
    fallback_dense: int
    bytes: int
    build_s: dict


def plan_layout(r0, m, c0, nn, adm, lr_idx, ranks):
    """Storage plan of the pool (element offsets).

    Admissible blocks ``lr_idx`` with ACA rank ``ranks`` (-1: not converged)
    are stored as U (3m x k) then V (k x 2n) unless they did not converge or
    their factors are not smaller than the six dense m x n operators; every
    other block is stored densely (six column-major m x n planes)."""
    lr_idx = np.asarray(lr_idx, dtype=np.int64)
    ranks = np.asarray(ranks, dtype=np.int64)
    mm, nl = m[lr_idx], nn[lr_idx]
    to_dense = (ranks < 0) | (ranks * (3 * mm + 2 * nl) >= 6 * mm * nl)
    keep = ~to_dense
    dense_idx = np.concatenate([np.where(~adm)[0], lr_idx[to_dense]]).astype(np.int64)
    d_size = 6 * m[dense_idx] * nn[dense_idx]
    d_base = np.concatenate([[0], np.cumsum(d_size)[:-1]]).astype(np.int64)
    off = int(d_size.sum())
    k = ranks[keep]
    u_size = 3 * mm[keep] * k
    v_size = 2 * nl[keep] * k
    l_ub = off + np.concatenate([[0], np.cumsum(u_size + v_size)[:-1]]).astype(np.int64)
    l_vb = l_ub + u_size
    total = off + int((u_size + v_size).sum())
    return dict(dense_idx=dense_idx, d_base=d_base, d_size=d_size, lr_idx=lr_idx[keep],
                l_ub=l_ub, l_vb=l_vb, k=k, keep=keep, fallback=int((ranks < 0).sum()),
                total=total, u_size=u_size, v_size=v_size)


def leaf_tables(tree):
    """(tiles, tile_of_row): the leaf clusters as (first row, rows) in cluster
    order, and the leaf index of every cluster-order row. Leaves tile the rows
    without gaps, so every block's rows are a run of whole leaves."""
    leaves = tree.leaves()
    lstart = tree.start[leaves].astype(np.int64)
    lsize = tree.size[leaves].astype(np.int64)
    if lstart[0] != 0 or np.any(lstart[1:] != lstart[:-1] + lsize[:-1]):
        raise ValueError("leaf clusters must tile the rows in order")
    tile_of_row = np.repeat(np.arange(len(leaves), dtype=np.int64), lsize)
    return np.ascontiguousarray(np.stack([lstart, lsize], axis=1)), tile_of_row


def leaf_segments(tiles, tile_of_row, r0, m):
    """(row offset in the block, rows) of each leaf of block rows [r0, r0+m)."""
    q0, q1 = int(tile_of_row[r0]), int(tile_of_row[r0 + m - 1])
    seg = tiles[q0:q1 + 1].copy()
    seg[:, 0] -= r0
    if seg[0, 0] != 0 or int(seg[:, 1].sum()) != m:
        raise ValueError("block rows are not a run of whole leaves")
    return seg


def pack_u(U, seg):
    """U (3m x k, rows o*m + i) in pool order: one panel per leaf segment
    (ro, ls) at offset 3*k*ro, plane-major inside the panel: element
    (o*m + i, l) at o*k*ls + l*ls + i - ro. A row tile then reads one
    contiguous panel, sharing no cache line with the tiles of the block's
    other leaves, and the rank columns of one row plane are contiguous (a
    product that skips the third plane skips a whole sub-panel)."""
    m3, k = U.shape
    m = m3 // 3
    U3 = U.reshape(3, m, k)
    out = np.empty(m3 * k, dtype=U.dtype)
    for ro, ls in seg:
        out[3 * k * ro:3 * k * (ro + ls)] = U3[:, ro:ro + ls, :].transpose(0, 2, 1).ravel()
    return out


def unpack_u(flat, m, k, seg):
    """Inverse of ``pack_u``: (3m x k)."""
    U = np.empty((3, m, k), dtype=flat.dtype)
    for ro, ls in seg:
        U[:, ro:ro + ls, :] = flat[3 * k * ro:3 * k * (ro + ls)].reshape(3, k, ls).transpose(0, 2, 1)
    return U.reshape(3 * m, k)


def reduce_rows(reduce, nrhs):
    """Phase-1b rows for ``nrhs`` right-hand sides: each reduction (t row, k,
    first partial row, units) split into chunks of max(1, 64 // nrhs) ranks,
    so that a wavefront sums about 64 elements whatever nrhs is. Row:
    (t row, k, first partial row, units, first rank, ranks)."""
    reduce = np.asarray(reduce, dtype=np.int64).reshape(-1, 4)
    if len(reduce) == 0:
        return np.zeros((0, 6), dtype=np.int64)
    lc = max(1, 64 // nrhs)
    k = reduce[:, 1]
    nch = -(-k // lc)
    rows = np.repeat(reduce, nch, axis=0)
    first = np.concatenate([[0], np.cumsum(nch)[:-1]]).astype(np.int64)
    l0 = (np.arange(int(nch.sum()), dtype=np.int64) - np.repeat(first, nch)) * lc
    lcnt = np.minimum(lc, rows[:, 1] - l0)
    return np.ascontiguousarray(np.column_stack([rows, l0, lcnt]).astype(np.int64))


def pack_units(groups):
    """Phase-1 units from low-rank blocks grouped by source columns:
    ``groups[(c0, n)]`` lists (V base, k, t row, first partial row) per
    block. Blocks of a group go into wavefronts by first fit in order of
    decreasing rank (capacity 64 lanes; a block of rank > 64 is alone and
    its kernel loops over 64-rank chunks); each wavefront then covers the
    group's 2n stacked columns in chunks of UNIT_COLS. Returns (units (Q, 8),
    members (M, 4)) as described in ``build_worklists``."""
    units, members = [], []
    for (bc0, bn) in sorted(groups):
        blocks = sorted(groups[(bc0, bn)], key=lambda b: -b[1])
        bins = []
        for blk in blocks:
            k = blk[1]
            for bin_ in bins:
                if k <= 64 and bin_[0] + k <= 64:
                    bin_[0] += k
                    bin_[1].append(blk)
                    break
            else:
                bins.append([k, [blk]])
        C = 2 * bn
        nu = -(-C // UNIT_COLS)
        for u in range(nu):                  # units reading the same columns are adjacent
            cb = u * UNIT_COLS
            for ktot, blks in bins:
                mem0, lane0 = len(members), 0
                for vb, k, to, po in blks:
                    members.append((vb, k, to if nu == 1 else po + u * k, lane0))
                    lane0 += k
                units.append((mem0, len(blks), cb, min(UNIT_COLS, C - cb), bn, bc0,
                              1 if nu == 1 else 0, ktot))
    return (np.ascontiguousarray(np.array(units, dtype=np.int64).reshape(-1, 8)),
            np.ascontiguousarray(np.array(members, dtype=np.int64).reshape(-1, 4)))


def xcd_chunks(g):
    """Workgroup ranges of the eight XCDs for a grid of g workgroups, as the
    kernels' xcd_order maps them: XCD x takes a contiguous range of q + 1
    (x < r) or q workgroups, q = g // 8, r = g % 8."""
    q, r = divmod(int(g), 8)
    sizes = [q + 1 if x < r else q for x in range(8)]
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def xcd_balanced_order(cost, per_group):
    """Order of work items (``per_group`` per workgroup) that keeps every
    XCD's contiguous range of work but puts the costliest items of each
    range first: workgroups are dispatched in order, so the cheap ones fill
    the tail instead of a long one running alone. The order changes no
    result (every item writes its own rows)."""
    cost = np.asarray(cost, dtype=np.float64)
    w = len(cost)
    g = -(-w // per_group)
    b = xcd_chunks(g) * per_group
    out = []
    for x in range(8):
        lo, hi = min(int(b[x]), w), min(int(b[x + 1]), w)
        idx = np.arange(lo, hi)
        out.append(idx[np.argsort(-cost[lo:hi], kind="stable")])
    return np.concatenate(out).astype(np.int64)


def mfs_steps(items, tile_ptr, tiles, n, nt, waves=1, order="row"):
    """Phase-2 step stream of the 16-RHS matrix-core product
    (``pq3d_hmat_matvec_mfs``): every item of every row tile cut into steps
    of 4 columns. Returns (steps (S, 4) int32, wptr (T*waves + 1,) int64,
    tiles (T, 2) int64 in launch order).

    Step record: pool element of the step's first column at the tile's
    first row (plane 0), column stride, plane stride, (columns - 1) << 28 |
    weight row, where weight rows are x1 rows [0, n), x2 rows [n, 2n) and t
    rows [2n, 2n + nt) of the FP32 weight array. A tile's steps keep the
    order of its items; ``waves`` contiguous ranges of (nearly) equal length
    per tile go to its wavefronts. ``order``: "row" launches the tiles in
    leaf (row) order, "cost" keeps the work lists' XCD-balanced order,
    "rowbal" gives each XCD one run of leaf-order tiles of equal cost
    (tiles padded with empty ones, no rows and no steps, to 8 equal runs)."""
    it = np.asarray(items, dtype=np.int64).reshape(-1, 6)
    tp = np.asarray(tile_ptr, dtype=np.int64)
    T = np.asarray(tiles, dtype=np.int64).reshape(-1, 2)
    ntiles = len(T)
    if waves not in MFS_WAVES:
        raise ValueError(f"waves must be one of {MFS_WAVES}")
    item_tile = np.repeat(np.arange(ntiles, dtype=np.int64), np.diff(tp))
    if len(item_tile) != len(it):
        raise ValueError("items and tile_ptr disagree")
    if order == "row":
        o = np.argsort(T[:, 0], kind="stable")
        rank = np.empty(ntiles, dtype=np.int64)
        rank[o] = np.arange(ntiles)
        T = T[o]
        item_tile = rank[item_tile]
        io = np.argsort(item_tile, kind="stable")
        it, item_tile = it[io], item_tile[io]
    elif order == "rowbal":
        # leaf order, cut into eight ranges of equal cost (columns), one per
        # XCD; launch slot b runs on XCD b mod 8 and the kernel reads tile
        # xcd_order(b), so range x is laid out at the xcd_order positions of
        # slots x, x + 8, ...; ranges shorter than the longest are padded
        # with empty tiles (no rows, no steps)
        o = np.argsort(T[:, 0], kind="stable")
        rank = np.empty(ntiles, dtype=np.int64)
        rank[o] = np.arange(ntiles)
        cost = np.bincount(rank[item_tile], weights=it[:, 4], minlength=ntiles)
        cuts = np.searchsorted(np.cumsum(cost), np.arange(1, 8) * cost.sum() / 8.0)
        bounds = np.concatenate([[0], cuts, [ntiles]])
        per = int(np.diff(bounds).max()) if ntiles else 0
        g = 8 * per
        q, r = divmod(g, 8)
        slot_pos = np.empty(g, dtype=np.int64)      # xcd_order(b) for every slot b
        b = np.arange(g)
        x, i = b % 8, b // 8
        slot_pos[:] = np.where(x < r, x * (q + 1) + i, r * (q + 1) + (x - r) * q + i)
        tile_at = np.full(g, -1, dtype=np.int64)    # position -> row-order tile (or -1: empty)
        for xx in range(8):
            rng = np.arange(bounds[xx], bounds[xx + 1])
            tile_at[slot_pos[xx + 8 * np.arange(len(rng))]] = rng
        Tn = np.zeros((g, 2), dtype=np.int64)
        Tn[tile_at >= 0] = T[o][tile_at[tile_at >= 0]]
        newpos = np.empty(ntiles, dtype=np.int64)
        newpos[tile_at[tile_at >= 0]] = np.where(tile_at >= 0)[0]
        item_tile = newpos[rank[item_tile]]
        io = np.argsort(item_tile, kind="stable")
        it, item_tile = it[io], item_tile[io]
        T, ntiles = Tn, g
    elif order != "cost":
        raise ValueError(f"order must be one of {MFS_ORDERS}")
    kind, base, ld, plane, cnt, xoff = (it[:, i] for i in range(6))
    nst = (cnt + 3) // 4
    rep = np.repeat(np.arange(len(it), dtype=np.int64), nst)
    first = np.concatenate([[0], np.cumsum(nst)[:-1]]).astype(np.int64)
    j = np.arange(len(rep), dtype=np.int64) - first[rep]
    rowbase = np.where(kind == 0, 0, np.where(kind == 2, n, 2 * n))[rep]
    sbase = base[rep] + 4 * j * ld[rep]
    scnt = np.minimum(4, cnt[rep] - 4 * j)
    srow = rowbase + xoff[rep] + 4 * j
    if len(rep):
        if sbase.max() >= 2 ** 32 or ld.max() >= 2 ** 31 or plane.max() >= 2 ** 31:
            raise ValueError("step record fields out of range (pool too large for the step stream)")
        if (srow + scnt).max() > 2 * n + nt or scnt.min() < 1 or 2 * n + nt >= 2 ** STEP_CNT_SHIFT:
            raise ValueError("step weight rows out of range")
    rec = np.empty((len(rep), 4), dtype=np.uint32)
    rec[:, 0] = sbase
    rec[:, 1] = ld[rep]
    rec[:, 2] = plane[rep]
    rec[:, 3] = ((scnt - 1).astype(np.uint32) << STEP_CNT_SHIFT) | srow.astype(np.uint32)
    per_tile = np.bincount(item_tile[rep], minlength=ntiles).astype(np.int64)
    tsp = np.concatenate([[0], np.cumsum(per_tile)]).astype(np.int64)
    wptr = np.empty(ntiles * waves + 1, dtype=np.int64)
    for g in range(waves):
        wptr[g:ntiles * waves:waves] = tsp[:-1] + (g * per_tile) // waves
    wptr[-1] = tsp[-1]
    return np.ascontiguousarray(rec.view(np.int32)), wptr, np.ascontiguousarray(T)


def vpanel_plan(c0, nn, lr_idx, ranks, l_vb, unit_cols=VP_UNIT_COLS, align=VP_ALIGN):
    """V panels of the matrix-core product (``pq3d_hmat_vpanel_pack``,
    ``pq3d_hmat_matvec_mfs``). The kept low-rank blocks ``lr_idx`` (ranks,
    V^T bases ``l_vb`` in the pool) are grouped by source cluster (c0, n);
    a group's V^T are stored side by side as one 2n x K panel (K = the sum of
    their ranks, blocks in ``lr_idx`` order), chunk-major: ranks
    [64 j, 64 j + kc) of a panel form one contiguous 2n x kc row-major chunk.
    Panels start at multiples of ``align`` floats, so with the default 128 B
    every full chunk (2n x 64 floats after the panel base) starts a cache
    line and its 4-column steps (1 KiB) do not straddle lines; the padding
    is never read. t rows are numbered panel by panel, so a chunk's t rows
    are contiguous.

    Returns dict with
      tab    (B, 6): packing rows (V^T base, k, n, panel base, K, first rank)
      units  (U, 8): phase-1 units (chunk element of the first column, kc,
             output row, first stacked column, columns, n, c0, direct), one
             per column range of at most ``unit_cols`` stacked columns of a
             chunk, in XCD-balanced cost order
      reduce (R, 4): (t row, kc, first partial row, units) of chunks split
             over several units
      toff   (B,):   first t row of every block in the panel numbering
      size, nt, npart: panel storage in floats (with the padding), t rows,
             partial rows"""
    lr = np.asarray(lr_idx, dtype=np.int64)
    k = np.asarray(ranks, dtype=np.int64)
    vb = np.asarray(l_vb, dtype=np.int64)
    bc0 = np.asarray(c0, dtype=np.int64)[lr]
    bn = np.asarray(nn, dtype=np.int64)[lr]
    order = np.lexsort((bn, bc0))          # stable: lr order inside a group
    tab = np.zeros((len(lr), 6), dtype=np.int64)
    toff = np.zeros(len(lr), dtype=np.int64)
    units, reds = [], []
    pbase = tp = poff = 0
    i = 0
    while i < len(order):
        j = i
        while j < len(order) and bc0[order[j]] == bc0[order[i]] and bn[order[j]] == bn[order[i]]:
            j += 1
        mem = order[i:j]
        n_, c0_ = int(bn[mem[0]]), int(bc0[mem[0]])
        K = int(k[mem].sum())
        koff = np.concatenate([[0], np.cumsum(k[mem])[:-1]]).astype(np.int64)
        tab[mem] = np.stack([vb[mem], k[mem], np.full(len(mem), n_), np.full(len(mem), pbase),
                             np.full(len(mem), K), koff], axis=1)
        toff[mem] = tp + koff
        C = 2 * n_
        nc = -(-C // unit_cols)
        for l0 in range(0, K, 64):
            kc = min(64, K - l0)
            cbase = pbase + C * l0
            for u in range(nc):
                cb = u * unit_cols
                cc = min(unit_cols, C - cb)
                out = tp + l0 if nc == 1 else poff + u * kc
                units.append((cbase + cb * kc, kc, out, cb, cc, n_, c0_, 1 if nc == 1 else 0))
            if nc > 1:
                reds.append((tp + l0, kc, poff, nc))
                poff += nc * kc
        pbase += -(-C * K // align) * align
        tp += K
        i = j
    U = np.array(units, dtype=np.int64).reshape(-1, 8)
    if len(U):
        U = U[xcd_balanced_order(U[:, 4] * U[:, 1], 1)]
    return dict(tab=np.ascontiguousarray(tab), units=np.ascontiguousarray(U),
                reduce=np.array(reds, dtype=np.int64).reshape(-1, 4), toff=toff,
                size=int(pbase), nt=int(tp), npart=int(poff))


def check_item_extents(items, tiles, tile_ptr, esize) -> None:
    """Raise unless every phase-2 item spans less than 2^31 bytes: the kernel
    addresses an item through one buffer resource with 32-bit byte offsets
    (column offset (count - 1) ld, plane offset 2 plane + rows)."""
    it = np.asarray(items, dtype=np.int64).reshape(-1, 6)
    tp = np.asarray(tile_ptr, dtype=np.int64)
    mt = np.repeat(np.asarray(tiles, dtype=np.int64).reshape(-1, 2)[:, 1], np.diff(tp))
    if len(it) != len(mt):
        raise ValueError("items and tile_ptr disagree")
    span = ((it[:, 4] - 1) * it[:, 2] + 2 * it[:, 3] + mt) * int(esize)
    if len(span) and int(span.max()) >= 2 ** 31:
        raise ValueError("a matvec item spans 2 GiB or more")


def build_worklists(tree, n, r0, m, c0, nn, dense_idx, d_base, lr_idx, l_ub, l_vb, k_keep):
    """Matvec work lists.

    tiles   (T, 2): (first row, rows) of every leaf cluster (<= 64 rows)
    items   (I, 6): (kind, base, ld, plane, count, xoff), sorted by tile;
            every item is three row planes (outputs y1, y2, y3) of ``count``
            columns, element (o, i, j) at base + o*plane + j*ld + i, times
            weights w[xoff + j]: kind 0 = strike operators of a dense chunk
            (w = v1), kind 2 = its dip operators (w = v2), kind 1 = low-rank
            chunk of U (w = t)
    tile_ptr (T+1,): CSR offsets of each tile's items
    units   (Q, 8): phase-1 units, one wavefront each: (first member row,
            members, first stacked column, columns, block n, block c0,
            direct, total rank). The members of a unit are low-rank blocks
            with the same source columns (c0, n), so x is the same for all
            of them; a unit of several members has total rank <= 64 and its
            lanes run over (member, rank). Units reading the same columns
            are adjacent.
    members (M, 4): (V^T base, k, output row, first lane) per unit member;
            V^T of a block is 2n x k row-major at its V base, U is in leaf
            panels (``pack_u``). Output rows are the block's t rows (direct)
            or its partial rows for this column chunk.
    reduce  (R, 4): phase-1b sums (t row, k, first partial row, units); the
            kernel reads them as ``reduce_rows`` for the call's nrhs
    nt, npart: rows of t and of the partial buffer
    """
    tiles, tile_of_row = leaf_tables(tree)
    lstart, lsize = tiles[:, 0], tiles[:, 1]
    if np.any(lsize > 64):
        raise ValueError("leaf clusters larger than 64 rows are not supported by the tile kernel")
    ntiles = len(tiles)
    items = []
    for bi, base in zip(np.asarray(dense_idx).tolist(), np.asarray(d_base).tolist()):
        br0, bm, bc0, bn = int(r0[bi]), int(m[bi]), int(c0[bi]), int(nn[bi])
        for q in range(tile_of_row[br0], tile_of_row[br0 + bm - 1] + 1):
            ro = int(lstart[q]) - br0
            for j0 in range(0, bn, DENSE_CHUNK):
                cnt = min(DENSE_CHUNK, bn - j0)
                items.append((q, 0, base + j0 * bm + ro, bm, bm * bn, cnt, bc0 + j0))
                items.append((q, 2, base + 3 * bm * bn + j0 * bm + ro, bm, bm * bn, cnt, bc0 + j0))
    reds = []
    toff = 0
    poff = 0
    groups = {}   # (c0, n) -> [(V base, k, t row, first partial row)]
    for bi, ub, vb, k in zip(np.asarray(lr_idx).tolist(), np.asarray(l_ub).tolist(),
                             np.asarray(l_vb).tolist(), np.asarray(k_keep).tolist()):
        br0, bm, bc0, bn = int(r0[bi]), int(m[bi]), int(c0[bi]), int(nn[bi])
        nu = -(-2 * bn // UNIT_COLS)
        groups.setdefault((bc0, bn), []).append((vb, k, toff, poff))
        if nu > 1:
            reds.append((toff, k, poff, nu))
            poff += nu * k
        for q in range(tile_of_row[br0], tile_of_row[br0 + bm - 1] + 1):
            ro, ls = int(lstart[q]) - br0, int(lsize[q])
            pb = ub + 3 * k * ro   # this leaf's panel of U (plane-major)
            for l0 in range(0, k, LR_CHUNK):
                cnt = min(LR_CHUNK, k - l0)
                items.append((q, 1, pb + l0 * ls, ls, k * ls, cnt, toff + l0))
        toff += k
    units, members = pack_units(groups)
    it = np.array(items, dtype=np.int64).reshape(-1, 7)
    # tiles in XCD-balanced order (see xcd_balanced_order): cost = columns
    cost = np.bincount(it[:, 0], weights=it[:, 5], minlength=ntiles)
    order = xcd_balanced_order(cost, 1)
    rank = np.empty(ntiles, dtype=np.int64)
    rank[order] = np.arange(ntiles)
    tiles = np.ascontiguousarray(tiles[order])
    it[:, 0] = rank[it[:, 0]]
    it = it[np.argsort(it[:, 0], kind="stable")]
    counts = np.bincount(it[:, 0], minlength=ntiles)
    tile_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    if len(units):
        # phase-1 units likewise (4 per workgroup): cost = columns x ranks
        uo = xcd_balanced_order(units[:, 3] * units[:, 7], 4)
        units = np.ascontiguousarray(units[uo])
    return dict(items=np.ascontiguousarray(it[:, 1:]), tile_ptr=tile_ptr,
                tiles=tiles,
                units=units, members=members,
                reduce=np.ascontiguousarray(np.array(reds, dtype=np.int64).reshape(-1, 4)),
                nt=toff, npart=poff)


class HMatrix:
    def __init__(self, nodelst, elelst, mu, lambda_, *, leaf=64, eta=2.0, eps=1e-4,
                 dtype=torch.float32, device="cuda", kmax=128, aca_budget_gb=24.0,
                 jud_ele_order=False, verbose=False, halfspace=False, aca_wide=4096):
        t_all = time.perf_counter()
        self.device = torch.device(device)
        # admissible blocks with m + n >= aca_wide take the wide ACA (each rank
        # step spread over many workgroups); the rest one workgroup each. Both
        # give the same factors bit for bit.
        self.aca_wide = int(aca_wide)
        self.dtype = dtype
        self.esize = 4 if dtype == torch.float32 else 8
        if dtype not in (torch.float32, torch.float64):
            raise ValueError("dtype must be float32 or float64")
        self.accumulate = "f64"      # one of ACCUMULATE
        self.mu, self.lambda_ = float(mu), float(lambda_)
        nodelst = np.asarray(nodelst, dtype=np.float64)
        elelst = np.asarray(elelst, dtype=np.int64)
        self.eleVec, self.xg = get_eleVec(nodelst, elelst, jud_ele_order)
        e = elelst - 1
        tri = np.concatenate([nodelst[e[:, 0]], nodelst[e[:, 1]], nodelst[e[:, 2]]], axis=1)
        self.n = n = tri.shape[0]
        tm = {}
        t0 = time.perf_counter()
        self.tree = build_cluster_tree_device(self.xg, leaf, self.device)
        self.part = build_block_partition_device(self.tree, eta, device=self.device)
        self.leaf_tiles, self.tile_of_row = leaf_tables(self.tree)
        tm["tree"] = time.perf_counter() - t0
        self.perm = self.tree.perm
        self.iperm = np.empty_like(self.perm)
        self.iperm[self.perm] = np.arange(n)
        dev = self.device
        self.perm_t = torch.from_numpy(self.perm).to(dev)
        self.iperm_t = torch.from_numpy(self.iperm).to(dev)
        self.recv = torch.from_numpy(np.ascontiguousarray(
            receivers(self.eleVec, self.xg)[self.perm])).to(dev)
        self.rec = torch.from_numpy(source_records(tri[self.perm])).to(dev)
        self.tri_perm = np.ascontiguousarray(tri[self.perm])
        # half space (free surface z = 0): second per-triangle record
        self.halfspace = bool(halfspace)
        self.hrec = (torch.from_numpy(hs_records(self.tri_perm)).to(dev) if self.halfspace
                     else None)

        t = self.tree
        P = self.part
        r0 = t.start[P.rows]; m = t.size[P.rows]
        c0 = t.start[P.cols]; nn = t.size[P.cols]
        adm = P.adm.copy()
        stream = _lib.stream_handle()

        # ---------------- far field: ACA (batched by workspace budget)
        t0 = time.perf_counter()
        lr_idx = np.where(adm)[0]
        ranks = np.full(len(lr_idx), -1, dtype=np.int64)
        kcap = np.minimum(kmax, np.minimum(3 * m[lr_idx], 2 * nn[lr_idx]))
        self._ws = {}
        ws_per = kcap * (3 * m[lr_idx] + 2 * nn[lr_idx]) * 8
        budget = aca_budget_gb * 1e9
        lr_results = []  # (block indices, ranks, wsU, wsV, wsoff tensor)
        a = 0
        while a < len(lr_idx):
            b = a
            acc = 0
            while b < len(lr_idx) and (acc + ws_per[b] <= budget or b == a):
                acc += ws_per[b]
                b += 1
            sel = np.arange(a, b)
            res = self._run_aca(lr_idx[sel], r0, m, c0, nn, kcap[sel], eps, stream)
            ranks[sel] = res[0]
            lr_results.append((sel, res))
            a = b
        tm["aca"] = time.perf_counter() - t0

        L = plan_layout(r0, m, c0, nn, adm, lr_idx, ranks)
        dense_idx, d_base, d_size = L["dense_idx"], L["d_base"], L["d_size"]
        lr_global, l_ub, l_vb, k_keep = L["lr_idx"], L["l_ub"], L["l_vb"], L["k"]
        keep, fallback, total = L["keep"], L["fallback"], L["total"]
        u_size, v_size = L["u_size"], L["v_size"]
        d_m, d_n = m[dense_idx], nn[dense_idx]
        self.pool = torch.zeros(max(total, 1), dtype=dtype, device=dev)

        # dense assembly
        t0 = time.perf_counter()
        dtab = np.stack([r0[dense_idx], d_m, c0[dense_idx], d_n, d_base,
                         np.zeros_like(d_base)], axis=1).astype(np.int64)
        self.dtab = torch.from_numpy(np.ascontiguousarray(dtab)).to(dev)
        _lib.check_width("dense_block", self.dtab)
        _lib.check_width("recv", self.recv)
        _lib.check_width("rec", self.rec)
        _lib.call("pq3d_hmat_dense_assemble", _lib.ptr(self.dtab), len(dense_idx),
                  _lib.ptr(self.recv), _lib.ptr(self.rec), _lib.ptr(self.hrec), self.mu,
                  self.lambda_, _lib.ptr(self.pool), self.esize, stream)
        _sync(dev)
        tm["dense"] = time.perf_counter() - t0

        # pack low-rank factors (per ACA batch)
        t0 = time.perf_counter()
        lr_global = lr_idx[keep]
        tiles_d = torch.from_numpy(self.leaf_tiles).to(dev)
        tor_d = torch.from_numpy(self.tile_of_row).to(dev)
        _lib.check_width("tile", tiles_d)
        ub_of = dict(zip(lr_global.tolist(), l_ub.tolist()))
        vb_of = dict(zip(lr_global.tolist(), l_vb.tolist()))
        for sel, res in lr_results:
            _, wsU, wsV, wsoff, blocks = res
            ltab = np.zeros((len(sel), 8), dtype=np.int64)
            for q, gb in enumerate(blocks.tolist()):
                kq = int(ranks[sel[q]])
                ok = keep[sel[q]]
                ltab[q] = (r0[gb], m[gb], c0[gb], nn[gb], ub_of.get(gb, 0), vb_of.get(gb, 0),
                           kq if ok else 0, 0)
            lt = torch.from_numpy(ltab).to(dev)
            _lib.check_width("lowrank_block", lt)
            _lib.check_width("aca_ws", wsoff)
            _lib.call("pq3d_hmat_aca_pack", _lib.ptr(lt), _lib.ptr(wsoff), len(sel),
                      _lib.ptr(tiles_d), _lib.ptr(tor_d), _lib.ptr(wsU), _lib.ptr(wsV),
                      _lib.ptr(self.pool), self.esize, stream)
        _sync(dev)
        del lr_results
        if dev.type == "cuda":
            torch.cuda.empty_cache()
        tm["pack"] = time.perf_counter() - t0

        # ---------------- matvec work lists
        t0 = time.perf_counter()
        self._build_worklists(r0, m, c0, nn, dense_idx, d_base, lr_global, l_ub, l_vb, k_keep)
        tm["worklists"] = time.perf_counter() - t0
        tm["total"] = time.perf_counter() - t_all
        self.stats = HStats(n=n, n_dense=len(dense_idx), n_lowrank=len(lr_global),
                            dense_entries=int(d_size.sum()),
                            lowrank_entries=int((u_size + v_size).sum()),
                            ranks_mean=float(k_keep.mean()) if len(k_keep) else 0.0,
                            ranks_max=int(k_keep.max()) if len(k_keep) else 0,
                            fallback_dense=fallback, bytes=total * self.esize, build_s=tm)
        # kept for checks
        self.blocks = dict(r0=r0, m=m, c0=c0, n=nn, dense_idx=dense_idx, d_base=d_base,
                           lr_idx=lr_global, ranks=k_keep, l_ub=l_ub, l_vb=l_vb)
        if verbose:
            print(self.stats, flush=True)

    # ------------------------------------------------------------------ ACA
    def _run_aca(self, blocks, r0, m, c0, nn, kcap, eps, stream):
        dev = self.device
        nb = len(blocks)
        mb, nbk = m[blocks], nn[blocks]
        uoff = np.concatenate([[0], np.cumsum(kcap * 3 * mb)[:-1]]).astype(np.int64)
        voff = np.concatenate([[0], np.cumsum(kcap * 2 * nbk)[:-1]]).astype(np.int64)
        foff = np.concatenate([[0], np.cumsum(3 * mb + 2 * nbk)[:-1]]).astype(np.int64)
        wsU = torch.empty(int((kcap * 3 * mb).sum()), dtype=torch.float64, device=dev)
        wsV = torch.empty(int((kcap * 2 * nbk).sum()), dtype=torch.float64, device=dev)
        flags = torch.empty(int((3 * mb + 2 * nbk).sum()), dtype=torch.uint8, device=dev)
        wsoff = torch.from_numpy(np.stack([uoff, voff, foff, np.zeros_like(uoff)],
                                          axis=1).astype(np.int64).copy()).to(dev)
        # per-block kcap travels in the table's k column
        tab = np.zeros((nb, 8), dtype=np.int64)
        tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3] = r0[blocks], mb, c0[blocks], nbk
        tab[:, 6] = kcap
        tab_t = torch.from_numpy(tab).to(dev)
        _lib.check_width("lowrank_block", tab_t)
        _lib.check_width("aca_ws", wsoff)
        rk = np.empty(nb, dtype=np.int64)
        wide = (mb + nbk) >= self.aca_wide
        for sel, entry in ((~wide, "pq3d_hmat_aca"), (wide, "pq3d_hmat_aca_wide")):
            q = np.where(sel)[0]
            if not len(q):
                continue
            qt = torch.from_numpy(q).to(dev)
            t_q, ws_q = tab_t[qt].contiguous(), wsoff[qt].contiguous()
            kq = int(kcap[q].max())
            rank = torch.empty(len(q), dtype=torch.int64, device=dev)
            args = [_lib.ptr(t_q), _lib.ptr(ws_q), len(q), _lib.ptr(self.recv),
                    _lib.ptr(self.rec), _lib.ptr(self.hrec), self.mu, self.lambda_, float(eps),
                    kq, _lib.ptr(wsU), _lib.ptr(wsV), _lib.ptr(flags), _lib.ptr(rank)]
            if entry == "pq3d_hmat_aca_wide":
                ist = torch.empty(8 * len(q) + 1, dtype=torch.int64, device=dev)
                dst = torch.empty(len(q) * (kq + 1), dtype=torch.float64, device=dev)
                args += [_lib.ptr(ist), _lib.ptr(dst)]
            _lib.call(entry, *args, stream)
            rk[q] = rank.cpu().numpy()
        return rk, wsU, wsV, wsoff, blocks

    # ------------------------------------------------------------ work lists
    def _build_worklists(self, r0, m, c0, nn, dense_idx, d_base, lr_idx, l_ub, l_vb, k_keep):
        W = build_worklists(self.tree, self.n, r0, m, c0, nn, dense_idx, d_base, lr_idx, l_ub,
                            l_vb, k_keep)
        check_item_extents(W["items"], W["tiles"], W["tile_ptr"], self.esize)
        dev = self.device
        self.items = torch.from_numpy(W["items"]).to(dev)
        self.tile_ptr = torch.from_numpy(W["tile_ptr"]).to(dev)
        self.tiles = torch.from_numpy(W["tiles"]).to(dev)
        self.ntiles = W["tiles"].shape[0]
        self.units = torch.from_numpy(W["units"]).to(dev)
        self.nunits = W["units"].shape[0]
        self.members = torch.from_numpy(W["members"]).to(dev)
        self.reduce = torch.from_numpy(W["reduce"]).to(dev)
        self.nred = W["reduce"].shape[0]
        self.nt, self.npart = W["nt"], W["npart"]
        self.min_tile_rows = int(W["tiles"][:, 1].min()) if len(W["tiles"]) else 0
        self._scratch = {}
        _lib.check_width("item", self.items)
        _lib.check_width("unit", self.units)
        _lib.check_width("member", self.members)
        _lib.check_width("tile", self.tiles)
        # algorithmic bytes of one matvec: stored operators + 2 input + 3 output vectors
        self.matvec_bytes = int(self.pool.numel()) * self.esize + 5 * 8 * self.n

    def _bufs(self, nrhs):
        """Per right-hand-side count: t and partial scratch, and the phase-1b
        rows (one wavefront per row)."""
        if nrhs not in self._scratch:
            red = reduce_rows(self.reduce.cpu().numpy(), nrhs)
            red_d = torch.from_numpy(red).to(self.device)
            _lib.check_width("reduce", red_d)
            self._scratch[nrhs] = (
                torch.zeros(max(self.nt, 1) * nrhs, dtype=torch.float64, device=self.device),
                torch.zeros(max(self.npart, 1) * nrhs, dtype=torch.float64, device=self.device),
                red_d)
            if nrhs == 16 and self.esize == 4:
                # created here so that graph capture finds them
                self._x32(nrhs)
                if isinstance(self, HMatrix) and self.min_tile_rows >= 4:
                    self._mfs_lists()
        return self._scratch[nrhs]

    # ----------------------------------------------------------------- apply
    def matvec(self, v1, v2, out=None, stream=None, nrhs=1, instance_major=False, planes=3):
        """(AdotV1, AdotV2, dsigma_K) = six-operator product, cluster order,
        FP64. With nrhs > 1, v1/v2 rows are interleaved by instance (element
        (cell, b) at cell*nrhs + b) and out has shape (3, n*nrhs), in the
        same layout or, with ``instance_major``, element (cell, b) at
        b*n + cell. ``planes=2`` skips the Bs and Bd operators (the
        reference's step with a fixed normal traction, src/QDsim.py:906-909):
        out[2] is then left as it was."""
        n = self.n
        if nrhs not in NRHS:
            raise ValueError(f"nrhs must be one of {NRHS}")
        if planes not in (2, 3):
            raise ValueError("planes must be 2 or 3")
        if out is None:
            out = torch.empty((3, n * nrhs), dtype=torch.float64, device=self.device)
        prefilled = v1 is None and v2 is None
        if prefilled:
            # the rounded inputs are already in the matrix-core weights
            # (pq3d_deficit_w32 into weights())
            if not (self.accumulate == "mfma" and nrhs == 16 and self.esize == 4 and
                    isinstance(self, HMatrix)):
                raise ValueError("v1 = v2 = None needs the 16-RHS matrix-core product")
        elif not (v1.dtype == torch.float64 and v2.dtype == torch.float64):
            raise ValueError("v1, v2 must be float64")
        elif not (v1.is_contiguous() and v2.is_contiguous()):
            raise ValueError("v1, v2 must be contiguous")
        elif v1.numel() != n * nrhs or v2.numel() != n * nrhs:
            raise ValueError("v1, v2 must hold n*nrhs values")
        if tuple(out.shape) != (3, n * nrhs) or not out.is_contiguous() or out.dtype != torch.float64:
            raise ValueError("out must be a contiguous (3, n*nrhs) float64 tensor")
        t, part, red = self._bufs(nrhs)
        if self.accumulate not in ACCUMULATE:
            raise ValueError(f"accumulate must be one of {ACCUMULATE}")
        mfma = self.accumulate.startswith("mfma") and nrhs == 16 and self.esize == 4
        if mfma and self.min_tile_rows < 4:
            raise ValueError("the matrix-core product needs tiles of at least 4 rows")
        y3 = _lib.ptr(out[2]) if planes == 3 else None
        args = [_lib.ptr(self.units), self.nunits, _lib.ptr(self.members), _lib.ptr(red),
                int(red.shape[0]), _lib.ptr(self.tiles), _lib.ptr(self.tile_ptr), self.ntiles,
                _lib.ptr(self.items), _lib.ptr(self.pool), self.esize, nrhs, _lib.ptr(v1),
                _lib.ptr(v2), _lib.ptr(t), _lib.ptr(part), _lib.ptr(out[0]), _lib.ptr(out[1]),
                y3, n if (instance_major and nrhs > 1) else 0]
        if mfma and self.accumulate == "mfma" and isinstance(self, HMatrix):
            steps, wptr, tl = self._mfs_lists()
            vp = self._vpanels()
            cfg = self._tuning()
            _lib.call("pq3d_hmat_matvec_mfs", _lib.ptr(vp["units"]), vp["nunits"],
                      _lib.ptr(vp["vp"]), _lib.ptr(vp["red"]), vp["nred"], _lib.ptr(steps),
                      _lib.ptr(wptr), _lib.ptr(tl), int(tl.shape[0]), _lib.ptr(self.pool),
                      _lib.ptr(v1), _lib.ptr(v2), _lib.ptr(self._x32(nrhs)), n, self.nt,
                      _lib.ptr(vp["part"]), _lib.ptr(out[0]), _lib.ptr(out[1]), y3,
                      n if (instance_major and nrhs > 1) else 0, int(cfg["p1_depth"]),
                      int(cfg["p1_waves"]), int(cfg["depth"]), int(cfg["waves"]),
                      _lib.stream_handle(stream))
        elif mfma:
            _lib.call("pq3d_hmat_matvec_mfma", *args, _lib.ptr(self._x32(nrhs)), n,
                      int(self.accumulate == "mfma_comp"), _lib.stream_handle(stream))
        else:
            _lib.call("pq3d_hmat_matvec", *args, _lib.stream_handle(stream))
        return out

    def weights(self) -> torch.Tensor:
        """The FP32 weight array of the 16-RHS matrix-core product (rows of
        16 floats: x1, x2, t, one zero row). Writing x1 and x2 there (e.g.
        with pq3d_deficit_w32) and calling ``matvec(None, None, nrhs=16)``
        skips the product's own input conversion."""
        return self._x32(16)

    def _x32(self, nrhs):
        """FP32 weights of the matrix-core products: the two inputs rounded
        (rows [0, 2n)), t (rows [2n, 2n + nt), written by phase 1 of the step
        stream product) and one row of zeros (read by the step stream for
        masked columns; nothing writes it)."""
        buf = getattr(self, "_x32_buf", None)
        rows = 2 * self.n + self.nt + 1
        if buf is None or buf.numel() < rows * nrhs:
            buf = torch.zeros(rows * nrhs, dtype=torch.float32, device=self.device)
            self._x32_buf = buf
        return buf

    def _tuning(self) -> dict:
        cfg = getattr(self, "tuning", None)
        return dict(MFS_DEFAULT) if cfg is None else cfg

    def _vpanels(self) -> dict:
        """The matrix-core product's V panels (built and packed on the device
        on first use, ``vpanel_plan``): panel storage, phase-1 units and
        phase-1b rows for 16 right-hand sides, partial rows, and the t-row
        numbering of every low-rank block."""
        vp = getattr(self, "_vp", None)
        if vp is None:
            B = self.blocks
            plan = vpanel_plan(B["c0"], B["n"], B["lr_idx"], B["ranks"], B["l_vb"])
            if plan["nt"] != self.nt:
                raise RuntimeError("V panels and work lists disagree on the ranks")
            dev = self.device
            vp = {"vp": torch.empty(max(plan["size"], 1), dtype=torch.float32, device=dev)}
            tab = torch.from_numpy(plan["tab"]).to(dev)
            _lib.check_width("vpanel", tab)
            _lib.call("pq3d_hmat_vpanel_pack", _lib.ptr(tab), len(plan["tab"]),
                      _lib.ptr(self.pool), _lib.ptr(vp["vp"]), _lib.stream_handle())
            vp["units"] = torch.from_numpy(plan["units"]).to(dev)
            _lib.check_width("punit", vp["units"])
            vp["nunits"] = len(plan["units"])
            red = reduce_rows(plan["reduce"], 16)
            vp["red"] = torch.from_numpy(red).to(dev)
            vp["nred"] = len(red)
            vp["part"] = torch.zeros(max(plan["npart"], 1) * 16, dtype=torch.float32, device=dev)
            vp["toff"] = plan["toff"]
            vp["bytes"] = plan["size"] * 4
            _sync(dev)
            self._vp = vp
        return vp

    def _mfs_lists(self, cfg=None):
        """Device step stream (steps, wptr, tiles) of the matrix-core product
        for the tuning ``cfg`` (default: the current one), built once per
        (tile order, wavefronts per tile). Its t rows follow the V panels'
        numbering."""
        cfg = self._tuning() if cfg is None else cfg
        cache = getattr(self, "_mfs_cache", None)
        if cache is None:
            cache = self._mfs_cache = {}
        key = (cfg["order"], int(cfg["waves"]))
        if key not in cache:
            host = getattr(self, "_wl_host", None)
            if host is None:
                items = self.items.cpu().numpy().copy()
                # U items address t by the work lists' numbering (blocks in
                # lr order); renumber to the panels'
                k = np.asarray(self.blocks["ranks"], dtype=np.int64)
                old = np.concatenate([[0], np.cumsum(k)[:-1]]).astype(np.int64)
                new = self._vpanels()["toff"]
                u = items[:, 0] == 1
                b = np.searchsorted(old, items[u, 5], side="right") - 1
                items[u, 5] = new[b] + (items[u, 5] - old[b])
                host = self._wl_host = (items, self.tile_ptr.cpu().numpy(),
                                        self.tiles.cpu().numpy())
            st, wp, tl = mfs_steps(*host, self.n, self.nt, waves=key[1], order=key[0])
            dev = self.device
            cache[key] = (torch.from_numpy(st).to(dev), torch.from_numpy(wp).to(dev),
                          torch.from_numpy(tl).to(dev))
        return cache[key]

    def tune(self, planes=2, reps=10, depths=MFS_DEPTHS, waves=MFS_WAVES, orders=MFS_ORDERS,
             p1_depths=MFP_DEPTHS, p1_waves=MFP_WAVES, seed=0) -> dict:
        """Autotuner of the 16-RHS matrix-core product, timed on this
        operator with HIP events (``reps`` products per candidate after two
        warm-up products, ``planes`` row planes, seeded inputs): first every
        phase-2 candidate (steps in flight per wavefront, wavefronts per row
        tile, tile launch order) with the default phase 1, then every
        phase-1 candidate (steps in flight, wavefronts per unit) with the
        chosen phase 2; the fastest is kept as ``tuning``. A choice changes
        only the FP32 summation order within a tile or unit. Returns
        {"choice", "planes", "candidates": [(config, ms)]}, also kept as
        ``tuning_report``."""
        if not isinstance(self, HMatrix) or self.esize != 4 or self.min_tile_rows < 4:
            raise ValueError("the matrix-core product needs an FP32 HMatrix with tiles of >= 4 rows")
        n, nb, dev = self.n, 16, self.device
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        v1 = torch.randn(n * nb, dtype=torch.float64, device=dev, generator=gen)
        v2 = torch.randn(n * nb, dtype=torch.float64, device=dev, generator=gen)
        out = torch.empty((3, n * nb), dtype=torch.float64, device=dev)
        acc0, cfg0 = self.accumulate, getattr(self, "tuning", None)
        self.accumulate = "mfma"
        results = []

        def run(cfg):
            self.tuning = dict(cfg)
            for _ in range(2):
                self.matvec(v1, v2, out=out, nrhs=nb, planes=planes)
            ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(reps)]
            for a, b in ev:
                a.record()
                self.matvec(v1, v2, out=out, nrhs=nb, planes=planes)
                b.record()
            torch.cuda.synchronize(dev)
            ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
            results.append((dict(cfg), ms))
            return ms

        try:
            base = dict(MFS_DEFAULT)
            p2 = [dict(base, order=o, waves=int(w), depth=int(d))
                  for o in orders for w in waves for d in depths]
            best = min(p2, key=run)
            p1 = [dict(best, p1_depth=int(d), p1_waves=int(w))
                  for w in p1_waves for d in p1_depths
                  if (int(d), int(w)) != (best["p1_depth"], best["p1_waves"])]
            for cfg in p1:
                run(cfg)
        finally:
            self.accumulate = acc0
            self.tuning = cfg0
        best = dict(min(results, key=lambda r: r[1])[0])
        self.tuning = best
        # only the chosen step stream stays on the device
        key = (best["order"], best["waves"])
        self._mfs_cache = {key: self._mfs_cache[key]}
        self.tuning_report = {"choice": best, "planes": planes, "candidates": results}
        return self.tuning_report

    def algorithmic_bytes(self, nrhs=1, planes=3):
        """Bytes one product must move at least: every stored operator
        element it applies, once (the third row plane of the dense and U
        storage only with ``planes=3``), plus two FP64 input and ``planes``
        FP64 output vectors per right-hand side."""
        B = self.blocks
        m, nn = np.asarray(B["m"]), np.asarray(B["n"])
        di = np.asarray(B["dense_idx"])
        li = np.asarray(B["lr_idx"])
        k = np.asarray(B["ranks"])
        dense = int((6 * m[di] * nn[di]).sum())
        u = int((3 * m[li] * k).sum())
        v = int((2 * nn[li] * k).sum())
        ops = v + (dense + u) * planes // 3
        return ops * self.esize + (2 + planes) * 8 * self.n * nrhs

    def to_cluster(self, x: torch.Tensor) -> torch.Tensor:
        """Original element order -> cluster order."""
        return x[..., self.perm_t]

    def from_cluster(self, x: torch.Tensor) -> torch.Tensor:
        return x[..., self.iperm_t]

    # ------------------------------------------------------------- storage
    FORMAT = "pq3d-hmatrix"
    VERSION = 2      # 2: plane-major U panels
    _BLOCK_KEYS = ("r0", "m", "c0", "n", "dense_idx", "d_base", "lr_idx", "ranks", "l_ub", "l_vb")

    def save(self, path, **meta) -> None:
        """Write the built operator (storage pool, work lists, permutation,
        block tables) to one safetensors file with a versioned header: the
        cache that replaces the reference's per-block pickles
        (src/Hmatrix.py:1240-1241, src/QDsim.py:199-218). ``meta`` (strings)
        is stored alongside, e.g. the model and build parameters."""
        from safetensors.torch import save_file
        cpu = lambda t: t.detach().to("cpu").contiguous()
        tens = {"pool": cpu(self.pool), "units": cpu(self.units), "members": cpu(self.members),
                "reduce": cpu(self.reduce), "tiles": cpu(self.tiles),
                "tile_ptr": cpu(self.tile_ptr), "items": cpu(self.items),
                "perm": torch.from_numpy(np.ascontiguousarray(self.perm, dtype=np.int64)),
                "leaf_tiles": torch.from_numpy(np.ascontiguousarray(self.leaf_tiles)),
                "tile_of_row": torch.from_numpy(np.ascontiguousarray(self.tile_of_row))}
        for k in self._BLOCK_KEYS:
            tens["block_" + k] = torch.from_numpy(np.ascontiguousarray(self.blocks[k],
                                                                       dtype=np.int64))
        st = self.stats
        header = {"format": self.FORMAT, "version": str(self.VERSION), "n": str(self.n),
                  "nt": str(self.nt), "npart": str(self.npart), "esize": str(self.esize),
                  "halfspace": str(int(self.halfspace)), "mu": repr(self.mu),
                  "lambda": repr(self.lambda_), "n_dense": str(st.n_dense),
                  "n_lowrank": str(st.n_lowrank), "dense_entries": str(st.dense_entries),
                  "lowrank_entries": str(st.lowrank_entries), "ranks_mean": repr(st.ranks_mean),
                  "ranks_max": str(st.ranks_max), "fallback_dense": str(st.fallback_dense)}
        cfg = getattr(self, "tuning", None)
        if cfg is not None:
            header["mfs_tuning"] = (f"{cfg['depth']},{cfg['waves']},{cfg['order']},"
                                    f"{cfg['p1_depth']},{cfg['p1_waves']}")
        header.update({"user_" + k: str(v) for k, v in meta.items()})
        save_file(tens, str(path), metadata=header)

    @classmethod
    def load(cls, path, device="cuda"):
        """An H-matrix written by ``save``, ready for ``matvec`` on
        ``device`` (nothing is re-assembled). Raises ValueError for another
        format or version."""
        from safetensors import safe_open
        with safe_open(str(path), framework="pt") as f:
            h = f.metadata() or {}
            if h.get("format") != cls.FORMAT or h.get("version") != str(cls.VERSION):
                raise ValueError(f"{path}: not a {cls.FORMAT} v{cls.VERSION} file")
            t = {k: f.get_tensor(k) for k in f.keys()}
        self = cls.__new__(cls)
        dev = torch.device(device)
        self.device = dev
        self.esize = int(h["esize"])
        self.dtype = torch.float32 if self.esize == 4 else torch.float64
        self.accumulate = "f64"
        self.n, self.nt, self.npart = int(h["n"]), int(h["nt"]), int(h["npart"])
        self.halfspace = h["halfspace"] == "1"
        self.mu, self.lambda_ = float(h["mu"]), float(h["lambda"])
        self.meta = {k[5:]: v for k, v in h.items() if k.startswith("user_")}
        if "mfs_tuning" in h:
            f5 = h["mfs_tuning"].split(",")
            try:
                cfg = {"depth": int(f5[0]), "waves": int(f5[1]), "order": f5[2],
                       "p1_depth": int(f5[3]), "p1_waves": int(f5[4])}
            except (IndexError, ValueError):
                raise ValueError(f"{path}: bad mfs_tuning {h['mfs_tuning']!r}") from None
            if (len(f5) != 5 or cfg["depth"] not in MFS_DEPTHS or cfg["waves"] not in MFS_WAVES
                    or cfg["order"] not in MFS_ORDERS or cfg["p1_depth"] not in MFP_DEPTHS
                    or cfg["p1_waves"] not in MFP_WAVES):
                raise ValueError(f"{path}: bad mfs_tuning {h['mfs_tuning']!r}")
            self.tuning = cfg
        self.pool = t["pool"].to(dev)
        for k in ("units", "members", "reduce", "tiles", "tile_ptr", "items"):
            setattr(self, k, t[k].to(dev))
        self.nunits, self.nred, self.ntiles = len(t["units"]), len(t["reduce"]), len(t["tiles"])
        self.min_tile_rows = int(t["tiles"][:, 1].min()) if self.ntiles else 0
        for name in ("unit", "member", "tile", "item"):
            _lib.check_width(name, getattr(self, name + "s"))
        check_item_extents(t["items"].numpy(), t["tiles"].numpy(), t["tile_ptr"].numpy(),
                           self.esize)
        self.perm = t["perm"].numpy()
        self.iperm = np.empty_like(self.perm)
        self.iperm[self.perm] = np.arange(len(self.perm))
        self.perm_t = torch.from_numpy(self.perm).to(dev)
        self.iperm_t = torch.from_numpy(self.iperm).to(dev)
        self.leaf_tiles = t["leaf_tiles"].numpy()
        self.tile_of_row = t["tile_of_row"].numpy()
        self.blocks = {k: t["block_" + k].numpy() for k in cls._BLOCK_KEYS}
        self.stats = HStats(n=self.n, n_dense=int(h["n_dense"]), n_lowrank=int(h["n_lowrank"]),
                            dense_entries=int(h["dense_entries"]),
                            lowrank_entries=int(h["lowrank_entries"]),
                            ranks_mean=float(h["ranks_mean"]), ranks_max=int(h["ranks_max"]),
                            fallback_dense=int(h["fallback_dense"]),
                            bytes=int(self.pool.numel()) * self.esize, build_s={})
        self._scratch = {}
        self.matvec_bytes = int(self.pool.numel()) * self.esize + 5 * 8 * self.n
        return self

    def lowrank_factors(self, q):
        """(U (3m x k), V (k x 2n)) of the q-th stored low-rank block, FP64 numpy."""
        B = self.blocks
        bi = B["lr_idx"][q]
        m, n, k = int(B["m"][bi]), int(B["n"][bi]), int(B["ranks"][q])
        ub, vb = int(B["l_ub"][q]), int(B["l_vb"][q])
        seg = leaf_segments(self.leaf_tiles, self.tile_of_row, int(B["r0"][bi]), m)
        U = unpack_u(self.pool[ub:ub + 3 * m * k].double().cpu().numpy(), m, k, seg)
        V = self.pool[vb:vb + 2 * n * k].double().cpu().numpy().reshape(2 * n, k).T
        return U, V

    @classmethod
    def from_mesh(cls, mesh: FaultMesh, mu, lambda_, **kw):
        return cls(mesh.nodelst, mesh.elelst, mu, lambda_, **kw)


class MultiHMatrix:
    """Independent H-matrices (one per mesh, same storage type) applied as one
    block-diagonal operator: every matvec phase is one launch over all of
    them. Instance b's cells occupy [off[b], off[b] + n_b) of the
    concatenated vectors, in that mesh's cluster order (``perms[b]``).
    Each instance's rows see exactly its own work items in its own order,
    so its product is bitwise that of its own HMatrix."""

    ragged = True

    def __init__(self, hms):
        if not hms:
            raise ValueError("need at least one H-matrix")
        if len({h.dtype for h in hms}) != 1 or len({str(h.device) for h in hms}) != 1:
            raise ValueError("all H-matrices must share storage type and device")
        self.parts = list(hms)
        self.device = hms[0].device
        self.dtype = hms[0].dtype
        self.esize = hms[0].esize
        self.accumulate = "f64"
        self.sizes = np.array([h.n for h in hms], dtype=np.int64)
        self.off = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n = int(self.off[-1])
        self.perms = [h.perm for h in hms]
        self.perm = np.concatenate([h.perm + o for h, o in zip(hms, self.off[:-1])])
        pool_off = np.cumsum([0] + [int(h.pool.numel()) for h in hms])
        t_off = np.cumsum([0] + [int(h.nt) for h in hms])
        p_off = np.cumsum([0] + [int(h.npart) for h in hms])
        units, mems, reds, tiles, items = [], [], [], [], []
        ptr = [np.zeros(1, dtype=np.int64)]
        nitems = nmem = 0
        for b, h in enumerate(hms):
            U = h.units.cpu().numpy().copy()
            Mb = h.members.cpu().numpy().copy()
            if len(U):
                # direct flag of every member's unit
                direct = np.zeros(len(Mb), dtype=bool)
                for mem0, cnt, dr in zip(U[:, 0], U[:, 1], U[:, 6]):
                    direct[mem0:mem0 + cnt] = dr != 0
                Mb[:, 0] += pool_off[b]                                 # V^T base
                Mb[:, 2] += np.where(direct, t_off[b], p_off[b])        # output row
                U[:, 0] += nmem                                         # first member row
                U[:, 5] += self.off[b]                                  # block c0
            nmem += len(Mb)
            units.append(U.reshape(-1, 8))
            mems.append(Mb.reshape(-1, 4))
            R = h.reduce.cpu().numpy().copy()
            if len(R):
                R[:, 0] += t_off[b]
                R[:, 2] += p_off[b]
            reds.append(R.reshape(-1, 4))
            T = h.tiles.cpu().numpy().copy()
            T[:, 0] += self.off[b]
            tiles.append(T)
            I = h.items.cpu().numpy().copy()
            if len(I):
                I[:, 1] += pool_off[b]
                I[:, 5] += np.where(I[:, 0] == 1, t_off[b], self.off[b])
            items.append(I.reshape(-1, 6))
            tp = h.tile_ptr.cpu().numpy()
            ptr.append(tp[1:] + nitems)
            nitems += int(tp[-1])
        dev = self.device
        cat = lambda xs, w: torch.from_numpy(
            np.ascontiguousarray(np.concatenate(xs, 0).reshape(-1, w).astype(np.int64))).to(dev)
        self.pool = torch.cat([h.pool for h in hms])
        self.units, self.nunits = cat(units, 8), int(sum(len(u) for u in units))
        self.members = cat(mems, 4)
        self.reduce, self.nred = cat(reds, 4), int(sum(len(r) for r in reds))
        self.tiles, self.ntiles = cat(tiles, 2), int(sum(len(t) for t in tiles))
        self.items = cat(items, 6)
        self.tile_ptr = torch.from_numpy(np.concatenate(ptr).astype(np.int64)).to(dev)
        self.nt, self.npart = int(t_off[-1]), int(p_off[-1])
        self.min_tile_rows = int(self.tiles[:, 1].min()) if self.ntiles else 0
        self._scratch = {}
        for name, tab in (("item", self.items), ("unit", self.units), ("member", self.members),
                          ("tile", self.tiles)):
            _lib.check_width(name, tab)
        self.matvec_bytes = int(self.pool.numel()) * self.esize + 5 * 8 * self.n

    _bufs = HMatrix._bufs
    matvec = HMatrix.matvec
    _x32 = HMatrix._x32
    _tuning = HMatrix._tuning

    def algorithmic_bytes(self, nrhs=1, planes=3):
        """``HMatrix.algorithmic_bytes`` summed over the instances (one
        right-hand side each)."""
        vec = (2 + planes) * 8
        return sum(h.algorithmic_bytes(1, planes) - vec * h.n for h in self.parts) + vec * self.n
