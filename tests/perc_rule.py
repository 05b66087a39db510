"""Cluster statistics of one phase from the connection rule of
src/percolation/percolation.cpp, by graph search (scipy), for the tests.

The reference (:75-114, :117-286) gives the interest voxels of the 2x2x2
cube with largest corner (i, j, k) their smallest label whenever one of the
six index permutations of (i, j, k) is an interest voxel.  The fixed point of
those sweeps labels every connected set of the graph "interest voxels that
share such a cube" with the label of its first voxel in raster order; the
statistics follow (:318-376).  This restates the rule with numpy and a
sparse-graph search, a different route from the oracle's sweeps."""
import itertools

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components


def active_cubes(f):
    """cube (x, y, z) is active when f at one of the six permutations of
    (x, y, z) is set (f cubic, boolean)"""
    return (f | f.transpose(0, 2, 1) | f.transpose(1, 0, 2) | f.transpose(1, 2, 0) | f.transpose(2, 1, 0)
            | f.transpose(2, 0, 1))


def components(f):
    """-> (number of sets, set index per voxel (-1 outside), raster index of
    each set's first voxel)"""
    N = f.shape[0]
    idx = np.arange(f.size, dtype=np.int64).reshape(f.shape)
    act = active_cubes(f)
    # corners of cube (x, y, z): (x - dx, y - dy, z - dz); padding with -1
    # (outside) at index -1 of every axis
    fp = np.zeros((N + 1,) * 3, dtype=bool)
    fp[1:, 1:, 1:] = f
    ip = np.full((N + 1,) * 3, -1, dtype=np.int64)
    ip[1:, 1:, 1:] = idx
    corners = []
    for dx, dy, dz in itertools.product((0, 1), repeat=3):
        sl = (slice(1 - dx, N + 1 - dx), slice(1 - dy, N + 1 - dy), slice(1 - dz, N + 1 - dz))
        corners.append((fp[sl] & act, ip[sl]))
    # star from the first interest corner of every active cube
    first = np.full(f.shape, -1, dtype=np.int64)
    for m, i in corners:
        take = m & (first < 0)
        first[take] = i[take]
    rows, cols = [], []
    for m, i in corners:
        sel = m & (first >= 0) & (i != first)
        rows.append(first[sel])
        cols.append(i[sel])
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    g = coo_matrix((np.ones(r.size, dtype=np.int8), (r, c)), shape=(f.size, f.size))
    _, lab = connected_components(g, directed=False)
    lab = np.where(f.reshape(-1), lab, -1)
    # renumber the sets by their first voxel in raster order
    flat = f.reshape(-1)
    uniq, first_pos = np.unique(lab[flat], return_index=True)
    firsts = np.flatnonzero(flat)[first_pos]
    order = np.argsort(firsts)
    remap = np.full(lab.max() + 2 if lab.size else 1, -1, dtype=np.int64)
    remap[uniq[order]] = np.arange(order.size)
    set_of = np.where(lab >= 0, remap[np.maximum(lab, 0)], -1)
    return order.size, set_of.reshape(f.shape), np.sort(firsts)


def expected(ids, nph, phase):
    """-> (C1, all_clusters, clusters_vol_frac, vol_intrstd_phs) of the
    reference's runPercolationAnalysis at its fixed point"""
    ids = np.asarray(ids, dtype=np.int32)
    f = ids == phase
    plus = nph + 1
    K, set_of, firsts = components(f)
    vol = float(f.sum())
    rank_of_voxel = np.cumsum(f.reshape(-1)) - 1
    all_clusters = (plus + rank_of_voxel[firsts]).astype(np.int32)
    sizes = np.bincount(set_of[f], minlength=K).astype(np.float64)
    fr = sizes / vol if K else sizes
    srt = np.sort(fr)[::-1]
    # rank = clusters with a strictly larger fraction
    rank = np.array([int(np.sum(srt > v)) for v in fr], dtype=np.int64)
    c1 = ids.copy()
    c1[f] = plus + rank[set_of[f]]
    return c1, all_clusters, srt, vol


def surface(ids, p1, p2):
    """voxels of p1 times their -x, -y, -z face neighbours of p2, -1 outside"""
    ids = np.asarray(ids)
    pad = np.full(tuple(s + 1 for s in ids.shape), -1, dtype=np.int64)
    pad[1:, 1:, 1:] = ids
    m = ids == p1
    c = 0
    for sl in ((slice(0, -1), slice(1, None), slice(1, None)), (slice(1, None), slice(0, -1), slice(1, None)),
               (slice(1, None), slice(1, None), slice(0, -1))):
        c += int(np.sum(m & (pad[sl] == p2)))
    return c


def snake(N):
    """one winding path of width one: in each x plane of even index, rows
    along z at even y joined at alternating ends; the planes joined at
    alternating corners.  Its first voxel in raster order is at the origin."""
    f = np.zeros((N, N, N), dtype=bool)
    rows = list(range(0, N, 2))
    flip = False
    for xi, x in enumerate(range(0, N, 2)):
        ys = rows if xi % 2 == 0 else rows[::-1]
        for r, y in enumerate(ys):
            f[x, y, :] = True
            if r + 1 < len(ys):
                y2 = ys[r + 1]
                end = N - 1 if (r % 2 == 0) != flip else 0
                f[x, min(y, y2) + 1, end] = True
        last_end = N - 1 if ((len(ys) - 1) % 2 == 0) != flip else 0
        if x + 2 < N:
            f[x + 1, ys[-1], last_end] = True
        flip = not flip if len(ys) % 2 == 0 else flip
    return f
