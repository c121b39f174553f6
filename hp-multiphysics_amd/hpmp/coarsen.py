"""Coarse meshes of an unstructured triangle mesh for h-multigrid, and the vertex
connections between neighbouring meshes.

The reference builds each coarser mesh by vertex removal (tri_mesh::coarsen,
tri_mesh/coarsen.cpp:10-247): along every edge boundary it keeps every other
boundary vertex, triangulates the coarsened boundary, and inserts the fine interior
vertices whose distance to the vertices already in the mesh is at least a factor
(1.5-2) times the fine mesh's local edge length.  `coarsen` follows that recipe
with its own steps:

* each boundary chain keeps its first vertex, every other vertex after it and its
  last vertex;
* interior vertices are taken greedily in index order under the same spacing rule
  (a uniform-grid hash finds the nearest kept vertex);
* the kept vertices are triangulated by Delaunay (scipy.spatial), and triangles whose
  centroid lies outside the fine mesh are dropped (holes and concave boundaries);
* every boundary side of the coarse mesh belongs to the edge boundary of the fine
  boundary side nearest to its midpoint, so the coarse boundaries keep the fine
  ids and essential masks.

`connect` forms the tables of tri_mesh::connect (tri_mesh/connect.cpp; ccnnct and
fcnnct, tri_mesh/tri_mesh.h:197-201): for every vertex of one mesh, the triangle of
the other mesh holding it and its barycentric weights (outside the other mesh: the
nearest triangle, weights clipped to be non-negative and renormalised).
The coarse meshes only shape the multigrid corrections; the converged solution is
that of the fine mesh whatever they are.
"""
from __future__ import annotations

import numpy as np

from .mesh import TriMesh, check_mesh


def build_mesh(pnts, tri, side_bdry, nebd, essential, names=None, ids=None) -> TriMesh:
    """A TriMesh from vertices and triangles (any orientation).  ``side_bdry(a, b)``
    returns the edge-boundary index (0 .. nebd-1) of the boundary side joining
    vertices a and b.  Sides are stored in the direction of their first triangle, so
    boundary sides run counter-clockwise (domain on the left); each edge boundary's
    sides are ordered into chains."""
    pnts = np.asarray(pnts, dtype=np.float64)
    tri = np.asarray(tri, dtype=np.int64).copy()
    P = pnts[tri]
    a2 = (P[:, 0, 0] - P[:, 2, 0]) * (P[:, 1, 1] - P[:, 0, 1]) - (P[:, 0, 1] - P[:, 2, 1]) * (P[:, 1, 0] - P[:, 0, 0])
    flip = a2 < 0
    tri[flip] = tri[flip][:, [0, 2, 1]]
    seg_of, seg_pnt = {}, []
    tri_seg = np.empty_like(tri)
    tri_sgn = np.empty_like(tri)
    for t in range(tri.shape[0]):
        for k in range(3):
            a, b = int(tri[t, (k + 1) % 3]), int(tri[t, (k + 2) % 3])
            key = (min(a, b), max(a, b))
            s = seg_of.get(key)
            if s is None:
                s = len(seg_pnt)
                seg_of[key] = s
                seg_pnt.append((a, b))
            tri_seg[t, k] = s
            tri_sgn[t, k] = 1 if seg_pnt[s][0] == a else -1
    seg_pnt = np.array(seg_pnt, dtype=np.int64).reshape(-1, 2)
    cnt = np.bincount(tri_seg.ravel(), minlength=seg_pnt.shape[0])
    bsides = np.flatnonzero(cnt == 1)
    groups = [[] for _ in range(nebd)]
    for s in bsides:
        groups[side_bdry(int(seg_pnt[s, 0]), int(seg_pnt[s, 1]))].append(int(s))
    ordered = []
    for g in groups:
        # chain the group's sides head to tail (starting where no side ends)
        nxt = {int(seg_pnt[s, 0]): s for s in g}
        ends = {int(seg_pnt[s, 1]) for s in g}
        starts = [s for s in g if int(seg_pnt[s, 0]) not in ends] or g[:1]
        seen, chain = set(), []
        for s0 in starts + g:
            s = s0
            while s is not None and s not in seen:
                seen.add(s)
                chain.append(s)
                s = nxt.get(int(seg_pnt[s, 1]))
        ordered.append(chain)
    ebdry_start = np.cumsum([0] + [len(c) for c in ordered])
    m = TriMesh(
        pnts=pnts,
        seg_pnt=seg_pnt.astype(np.int32),
        tri_pnt=tri.astype(np.int32),
        tri_seg=tri_seg.astype(np.int32),
        tri_sgn=tri_sgn.astype(np.int32),
        ebdry_start=ebdry_start.astype(np.int32),
        ebdry_seg=np.array([s for c in ordered for s in c], dtype=np.int32),
        ebdry_essential=np.asarray(essential, dtype=np.int32).reshape(nebd),
        ebdry_names=list(names) if names is not None else [],
        ebdry_ids=list(ids) if ids is not None else [],
    )
    check_mesh(m)
    return m


def _barycentric(P, x):
    """Barycentric weights of points x (n, 2) in triangles P (n, 3, 2)."""
    d = (P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1])
    w1 = ((x[:, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 2, 0] - P[:, 0, 0]) * (x[:, 1] - P[:, 0, 1])) / d
    w2 = ((P[:, 1, 0] - P[:, 0, 0]) * (x[:, 1] - P[:, 0, 1]) - (x[:, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1])) / d
    return np.stack([1.0 - w1 - w2, w1, w2], axis=1)


def locate(mesh: TriMesh, x, k: int = 16):
    """For every point of x (n, 2): the triangle of ``mesh`` holding it (the candidate,
    among the k triangles with the nearest centroids, with the largest smallest
    barycentric weight; ties to the lowest index), its weights, and whether it is
    inside (smallest weight >= -1e-12)."""
    from scipy.spatial import cKDTree
    x = np.asarray(x, dtype=np.float64)
    cent = mesh.pnts[mesh.tri_pnt].mean(axis=1)
    kk = min(k, mesh.ntri)
    _, cand = cKDTree(cent).query(x, k=kk)
    cand = np.sort(np.asarray(cand).reshape(x.shape[0], kk), axis=1)
    best_t = np.full(x.shape[0], -1, dtype=np.int64)
    best_m = np.full(x.shape[0], -np.inf)
    best_w = np.zeros((x.shape[0], 3))
    for col in range(kk):
        t = cand[:, col]
        w = _barycentric(mesh.pnts[mesh.tri_pnt[t]], x)
        mn = w.min(axis=1)
        better = mn > best_m + 1e-14
        best_t = np.where(better, t, best_t)
        best_m = np.where(better, mn, best_m)
        best_w = np.where(better[:, None], w, best_w)
    return best_t, best_w, best_m >= -1e-12


def coarsen(fine: TriMesh, factor: float = 1.6) -> TriMesh:
    """The next coarser mesh of ``fine`` for h-multigrid (see the module docstring)."""
    from scipy.spatial import Delaunay
    n = fine.npnt
    # local length scale: the shortest side at each vertex
    L = np.linalg.norm(fine.pnts[fine.seg_pnt[:, 0]] - fine.pnts[fine.seg_pnt[:, 1]], axis=1)
    h = np.full(n, np.inf)
    np.minimum.at(h, fine.seg_pnt[:, 0], L)
    np.minimum.at(h, fine.seg_pnt[:, 1], L)
    # boundary chains: first vertex, every other vertex, last vertex
    keep = []
    on_bdry = fine.boundary_vertex_mask()
    for b in range(fine.nebd):
        sides = fine.ebdry_seg[fine.ebdry_start[b]:fine.ebdry_start[b + 1]]
        if sides.size == 0:
            continue
        verts = [int(fine.seg_pnt[sides[0], 0])] + [int(fine.seg_pnt[s, 1]) for s in sides]
        chain = verts[0::2]
        if chain[-1] != verts[-1]:
            chain.append(verts[-1])
        keep.extend(chain)
    keep = list(dict.fromkeys(keep))
    # interior vertices by the spacing rule, in index order
    cell = max(float(np.max(h[np.isfinite(h)])) * factor, 1e-300)
    grid = {}

    def cell_of(p):
        return int(np.floor(p[0] / cell)), int(np.floor(p[1] / cell))

    for v in keep:
        grid.setdefault(cell_of(fine.pnts[v]), []).append(v)
    for v in range(n):
        if on_bdry[v]:
            continue
        p = fine.pnts[v]
        ci, cj = cell_of(p)
        r2 = (factor * h[v]) ** 2
        ok = True
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                for u in grid.get((ci + di, cj + dj), ()):
                    if np.sum((fine.pnts[u] - p) ** 2) < r2:
                        ok = False
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            keep.append(v)
            grid.setdefault((ci, cj), []).append(v)
    keep = np.array(keep, dtype=np.int64)
    pts = fine.pnts[keep]
    tri = Delaunay(pts).simplices
    # drop the triangles outside the fine domain, and the slivers Delaunay leaves on a
    # straight boundary
    _, _, inside = locate(fine, pts[tri].mean(axis=1))
    P = pts[tri]
    a2 = np.abs((P[:, 0, 0] - P[:, 2, 0]) * (P[:, 1, 1] - P[:, 0, 1]) - (P[:, 0, 1] - P[:, 2, 1]) * (P[:, 1, 0] - P[:, 0, 0]))
    tri = tri[inside & (a2 > 1e-12 * np.median(a2))]
    used = np.unique(tri)
    remap = np.full(pts.shape[0], -1, dtype=np.int64)
    remap[used] = np.arange(used.size)
    pts = pts[used]
    tri = remap[tri]
    # boundary sides take the edge boundary of the nearest fine boundary side
    bs = fine.ebdry_seg
    bidx = np.repeat(np.arange(fine.nebd), np.diff(fine.ebdry_start))
    fmid = 0.5 * (fine.pnts[fine.seg_pnt[bs, 0]] + fine.pnts[fine.seg_pnt[bs, 1]])
    from scipy.spatial import cKDTree
    ftree = cKDTree(fmid)

    def side_bdry(a, b):
        _, j = ftree.query(0.5 * (pts[a] + pts[b]))
        return int(bidx[j])

    return build_mesh(pts, tri, side_bdry, fine.nebd, fine.ebdry_essential,
                      names=fine.ebdry_names, ids=fine.ebdry_ids)


def connect(fine: TriMesh, coarse: TriMesh):
    """(cc_tri, cc_wt, fc_tri, fc_wt): the coarse triangle and weights of every fine
    vertex, the fine triangle and weights of every coarse vertex (the layout
    hpg_link_fine takes).  Points outside the other mesh get the nearest triangle with
    the weights clipped to >= 0 and renormalised."""
    def table(src_pts, dst):
        t, w, inside = locate(dst, src_pts)
        w = np.where(inside[:, None], w, np.clip(w, 0.0, None))
        w = w / w.sum(axis=1, keepdims=True)
        return t.astype(np.int32), w

    cc_tri, cc_wt = table(fine.pnts, coarse)
    fc_tri, fc_wt = table(coarse.pnts, fine)
    return cc_tri, cc_wt, fc_tri, fc_wt


def hierarchy(fine: TriMesh, ngrid: int, factor: float = 1.6, min_tri: int = 8):
    """Meshes and connection tables for up to ``ngrid`` levels below and including
    ``fine`` (stops when a mesh would have fewer than ``min_tri`` triangles).  Returns
    ``(meshes, links)`` as hpmp.multigrid.square_hierarchy does."""
    meshes, links = [fine], []
    while len(meshes) < ngrid:
        c = coarsen(meshes[-1], factor)
        if c.ntri < min_tri or c.ntri >= meshes[-1].ntri:
            break
        links.append(connect(meshes[-1], c))
        meshes.append(c)
    return meshes, links
