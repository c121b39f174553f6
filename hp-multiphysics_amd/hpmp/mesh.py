"""Structured triangle meshes in the reference's tri_mesh layout.

The reference stores a mesh as ``pnts(npnt,2)``, ``seg(nseg).pnt(2)`` and
``tri(ntri).pnt(3)/seg(3)/sgn(3)`` (tri_mesh/tri_mesh.h:34-98).  Side ``j`` of a
triangle runs from vertex ``(j+1)%3`` to vertex ``(j+2)%3`` and ``sgn(j)`` is +1
when the segment is stored in that direction (tri_mesh/input.cpp:363-381).
Boundary segments are stored counter-clockwise (domain on the left), which is
what the edge-boundary normal ``(dy, -dx)`` in tri_hp/hp_boundary.cpp:987-992
assumes.

``unit_square(n)`` splits the unit square into n x n squares, each cut by the
diagonal from its lower-left to its upper-right corner (the mesh that repeated
``tri_mesh -r`` refinement of ``square1.grd`` produces), optionally jittering the
interior vertices by U(-j*h, j*h) per coordinate with a fixed seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class TriMesh:
    pnts: np.ndarray          # (npnt, 2) float64
    seg_pnt: np.ndarray       # (nseg, 2) int32
    tri_pnt: np.ndarray       # (ntri, 3) int32
    tri_seg: np.ndarray       # (ntri, 3) int32
    tri_sgn: np.ndarray       # (ntri, 3) int32
    ebdry_start: np.ndarray   # (nebd+1,) int32, offsets into ebdry_seg
    ebdry_seg: np.ndarray     # ordered boundary segments, counter-clockwise
    ebdry_essential: np.ndarray  # (nebd,) bit mask of essential variables
    ebdry_names: list = field(default_factory=list)
    ebdry_ids: list = field(default_factory=list)   # boundary idnum of each edge boundary (.grd files)
    vbdry: list = field(default_factory=list)       # (idnum, vertex) vertex boundaries (.grd files)

    @property
    def npnt(self) -> int:
        return int(self.pnts.shape[0])

    @property
    def nseg(self) -> int:
        return int(self.seg_pnt.shape[0])

    @property
    def ntri(self) -> int:
        return int(self.tri_pnt.shape[0])

    @property
    def nebd(self) -> int:
        return int(self.ebdry_start.shape[0] - 1)

    def boundary_vertex_mask(self) -> np.ndarray:
        m = np.zeros(self.npnt, dtype=bool)
        m[self.seg_pnt[self.ebdry_seg].ravel()] = True
        return m

    def boundary_segment_mask(self) -> np.ndarray:
        m = np.zeros(self.nseg, dtype=bool)
        m[self.ebdry_seg] = True
        return m

    def area2(self) -> np.ndarray:
        """Twice the signed area of every triangle (tri_mesh/tprims.cpp:66-80)."""
        p0, p1, p2 = (self.pnts[self.tri_pnt[:, k]] for k in range(3))
        d1 = p0 - p2
        d2 = p1 - p0
        return d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0]


def unit_square(n: int, jitter: float = 0.0, seed: int = 0, essential_mask: int = 0b011) -> TriMesh:
    n = int(n)
    N1 = n + 1
    ii, jj = np.meshgrid(np.arange(N1), np.arange(N1), indexing="xy")
    pnts = np.stack([ii.ravel() / n, jj.ravel() / n], axis=1).astype(np.float64)
    if jitter:
        rng = np.random.default_rng(seed)
        h = 1.0 / n
        d = rng.uniform(-jitter * h, jitter * h, size=pnts.shape)
        interior = (ii.ravel() > 0) & (ii.ravel() < n) & (jj.ravel() > 0) & (jj.ravel() < n)
        pnts[interior] += d[interior]

    def vid(i, j):
        return j * N1 + i

    # segment numbering: per square (row-major) H(i,j), V(i,j), D(i,j); then the top
    # row H(i,n) and the right column V(n,j)
    sq_i, sq_j = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    sq_i = sq_i.ravel()
    sq_j = sq_j.ravel()
    sq = sq_j * n + sq_i
    nseg = 3 * n * n + 2 * n
    seg_pnt = np.zeros((nseg, 2), dtype=np.int64)
    seg_pnt[3 * sq + 0] = np.stack([vid(sq_i, sq_j), vid(sq_i + 1, sq_j)], 1)        # H(i,j)
    seg_pnt[3 * sq + 1] = np.stack([vid(sq_i, sq_j), vid(sq_i, sq_j + 1)], 1)        # V(i,j)
    seg_pnt[3 * sq + 2] = np.stack([vid(sq_i, sq_j), vid(sq_i + 1, sq_j + 1)], 1)    # D(i,j)
    top = 3 * n * n + np.arange(n)
    right = 3 * n * n + n + np.arange(n)
    seg_pnt[top] = np.stack([vid(np.arange(n) + 1, n), vid(np.arange(n), n)], 1)     # H(i,n) reversed
    seg_pnt[right] = np.stack([vid(n, np.arange(n)), vid(n, np.arange(n) + 1)], 1)   # V(n,j)
    left_sq = sq[sq_i == 0]
    seg_pnt[3 * left_sq + 1] = seg_pnt[3 * left_sq + 1][:, ::-1]                     # V(0,j) reversed

    def H(i, j):
        return np.where(j < n, 3 * (np.minimum(j, n - 1) * n + i) + 0, 3 * n * n + i)

    def V(i, j):
        return np.where(i < n, 3 * (j * n + np.minimum(i, n - 1)) + 1, 3 * n * n + n + j)

    def D(i, j):
        return 3 * (j * n + i) + 2

    a = vid(sq_i, sq_j)
    b = vid(sq_i + 1, sq_j)
    c = vid(sq_i + 1, sq_j + 1)
    d = vid(sq_i, sq_j + 1)
    t1 = np.stack([a, b, c], 1)
    t2 = np.stack([a, c, d], 1)
    s1 = np.stack([V(sq_i + 1, sq_j), D(sq_i, sq_j), H(sq_i, sq_j)], 1)
    s2 = np.stack([H(sq_i, sq_j + 1), V(sq_i, sq_j), D(sq_i, sq_j)], 1)
    tri_pnt = np.empty((2 * n * n, 3), dtype=np.int64)
    tri_seg = np.empty_like(tri_pnt)
    tri_pnt[0::2] = t1
    tri_pnt[1::2] = t2
    tri_seg[0::2] = s1
    tri_seg[1::2] = s2
    tri_sgn = np.empty_like(tri_pnt)
    for k in range(3):
        tri_sgn[:, k] = np.where(seg_pnt[tri_seg[:, k], 0] == tri_pnt[:, (k + 1) % 3], 1, -1)

    bottom = H(np.arange(n), np.zeros(n, dtype=np.int64))
    rgt = V(np.full(n, n), np.arange(n))
    tp = H(np.arange(n)[::-1], np.full(n, n))
    lft = V(np.zeros(n, dtype=np.int64), np.arange(n)[::-1])
    ebdry_seg = np.concatenate([bottom, rgt, tp, lft])
    ebdry_start = np.array([0, n, 2 * n, 3 * n, 4 * n])
    ebdry_ess = np.full(4, essential_mask)
    m = TriMesh(
        pnts=pnts,
        seg_pnt=seg_pnt.astype(np.int32),
        tri_pnt=tri_pnt.astype(np.int32),
        tri_seg=tri_seg.astype(np.int32),
        tri_sgn=tri_sgn.astype(np.int32),
        ebdry_start=ebdry_start.astype(np.int32),
        ebdry_seg=ebdry_seg.astype(np.int32),
        ebdry_essential=ebdry_ess.astype(np.int32),
        ebdry_names=["bottom", "right", "top", "left"],
    )
    check_mesh(m)
    return m


def coarsen_square(fine: TriMesh, n_fine: int) -> TriMesh:
    """The next coarser multigrid mesh of ``unit_square(n_fine, ...)``: the n/2 x n/2
    structured mesh whose vertex (i, j) is the fine vertex (2i, 2j), so the jittered
    positions are inherited and the coarse vertices are a subset of the fine ones
    (the reference coarsens its meshes by vertex removal, tri_mesh/blocks.cpp:1245-1260)."""
    if n_fine % 2:
        raise ValueError("coarsening needs an even number of squares")
    nc = n_fine // 2
    c = unit_square(nc, 0.0, 0, essential_mask=int(fine.ebdry_essential[0]))
    N1f, N1c = n_fine + 1, nc + 1
    jj, ii = np.divmod(np.arange(N1c * N1c), N1c)
    c.pnts = fine.pnts[(2 * jj) * N1f + 2 * ii].copy()
    c.ebdry_essential = fine.ebdry_essential.copy()
    check_mesh(c)
    return c


def _locate(points: np.ndarray, mesh: TriMesh, n: int):
    """Triangle of the structured n x n ``mesh`` holding each point and its barycentric
    weights (the triangle with the largest smallest weight among the two triangles of
    the nominal cell and of its 8 neighbours; ties go to the lowest triangle index)."""
    x = np.clip((points[:, 0] * n).astype(np.int64), 0, n - 1)
    y = np.clip((points[:, 1] * n).astype(np.int64), 0, n - 1)
    best_t = np.full(points.shape[0], -1, dtype=np.int64)
    best_m = np.full(points.shape[0], -np.inf)
    best_w = np.zeros((points.shape[0], 3))
    cand = []
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            cx, cy = x + di, y + dj
            ok = (cx >= 0) & (cx < n) & (cy >= 0) & (cy < n)
            sq = np.where(ok, cy * n + cx, 0)
            for k in (0, 1):
                cand.append((np.where(ok, 2 * sq + k, -1)))
    cand = np.stack(cand, 1)
    cand_sorted = np.sort(np.where(cand < 0, np.iinfo(np.int64).max, cand), axis=1)
    for col in range(cand_sorted.shape[1]):
        t = cand_sorted[:, col]
        ok = t != np.iinfo(np.int64).max
        tt = np.where(ok, t, 0)
        P = mesh.pnts[mesh.tri_pnt[tt]]
        d = (P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1])
        w1 = ((points[:, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 2, 0] - P[:, 0, 0]) * (points[:, 1] - P[:, 0, 1])) / d
        w2 = ((P[:, 1, 0] - P[:, 0, 0]) * (points[:, 1] - P[:, 0, 1]) - (points[:, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1])) / d
        w0 = 1.0 - w1 - w2
        mn = np.minimum(w0, np.minimum(w1, w2))
        better = ok & (mn > best_m + 1e-14)
        best_t = np.where(better, t, best_t)
        best_m = np.where(better, mn, best_m)
        best_w = np.where(better[:, None], np.stack([w0, w1, w2], 1), best_w)
    assert np.all(best_t >= 0)
    return best_t.astype(np.int32), best_w


def connect(fine: TriMesh, n_fine: int, coarse: TriMesh):
    """Vertex connections between a mesh and its coarse neighbour (tri_mesh.h:197-201):
    ccnnct = (coarse triangle, weights) of every fine vertex, fcnnct = (fine triangle,
    weights) of every coarse vertex."""
    cc_tri, cc_wt = _locate(fine.pnts, coarse, n_fine // 2)
    fc_tri, fc_wt = _locate(coarse.pnts, fine, n_fine)
    return cc_tri, cc_wt, fc_tri, fc_wt


def _require(ok, what: str) -> None:
    if not bool(np.all(ok)):
        raise ValueError(f"inconsistent mesh: {what}")


def check_mesh(m: TriMesh) -> None:
    """Structural invariants every mesh handed to the solver must satisfy; raises
    ValueError naming the first one that fails (the C-ABI checks index ranges only)."""
    _require(m.area2() > 0.0, "triangles must be counter-clockwise")
    # every segment used by one (boundary) or two (interior) triangles, opposite signs
    cnt = np.bincount(m.tri_seg.ravel(), minlength=m.nseg)
    _require((cnt == 1) | (cnt == 2), "a segment is used by no triangle or by more than two")
    bmask = m.boundary_segment_mask()
    _require(cnt[bmask] == 1, "a boundary segment is shared by two triangles")
    _require(cnt[~bmask] == 2, "an interior segment belongs to one triangle only")
    for k in range(3):
        s = m.tri_seg[:, k]
        a = m.tri_pnt[:, (k + 1) % 3]
        b = m.tri_pnt[:, (k + 2) % 3]
        fwd = (m.seg_pnt[s, 0] == a) & (m.seg_pnt[s, 1] == b)
        rev = (m.seg_pnt[s, 0] == b) & (m.seg_pnt[s, 1] == a)
        _require(fwd | rev, f"side {k} of a triangle does not join its vertices")
        _require(np.where(fwd, 1, -1) == m.tri_sgn[:, k], f"tri_sgn of side {k} disagrees with seg_pnt")
    # boundary segments owned with sgn = +1 (domain on the left)
    owner_sgn = np.zeros(m.nseg, dtype=np.int64)
    for k in range(3):
        np.add.at(owner_sgn, m.tri_seg[:, k], m.tri_sgn[:, k])
    _require(owner_sgn[bmask] == 1, "a boundary segment is not stored counter-clockwise")
    _require(owner_sgn[~bmask] == 0, "an interior segment has equal signs in both triangles")
