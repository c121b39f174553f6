"""Tetrahedral meshes in the reference's conventions (tet_mesh/tet_mesh.h:127-137).

Local numbering follows tet_mesh/setup.cpp: edge i of a tet joins local vertices
(2,3), (3,1), (2,1), (1,0), (2,0), (3,0) (:230-235), with sgn(i) = -1 when the
edge's stored first vertex is the tet's second one (:272-275); face i is opposite
vertex i and ordered (1,3,2), (0,2,3), (0,3,1), (0,1,2) (:305-308), which makes
dx/dr x dx/ds of the face map point out of the tet; every tet is positively
oriented (vertices 2 and 3 swapped otherwise, :77-81).
"""
from __future__ import annotations

from dataclasses import dataclass
from itertools import permutations

import numpy as np

VS = np.array([[2, 3], [3, 1], [2, 1], [1, 0], [2, 0], [3, 0]], dtype=np.int64)
VF = np.array([[1, 3, 2], [0, 2, 3], [0, 3, 1], [0, 1, 2]], dtype=np.int64)


@dataclass
class TetMesh:
    pnts: np.ndarray     # (npnt, 3) float64
    seg_pnt: np.ndarray  # (nseg, 2) int32
    tet_pnt: np.ndarray  # (ntet, 4) int32
    tet_seg: np.ndarray  # (ntet, 6) int32
    tet_sgn: np.ndarray  # (ntet, 6) int32
    bf_pnt: np.ndarray   # (nbf, 3) int32, outward ordering
    bf_seg: np.ndarray   # (nbf, 3) int32, edge opposite face vertex j
    bf_sgn: np.ndarray   # (nbf, 3) int32

    @property
    def npnt(self) -> int:
        return self.pnts.shape[0]

    @property
    def nseg(self) -> int:
        return self.seg_pnt.shape[0]

    @property
    def ntet(self) -> int:
        return self.tet_pnt.shape[0]

    @property
    def nbf(self) -> int:
        return self.bf_pnt.shape[0]

    def boundary_vertex_mask(self) -> np.ndarray:
        m = np.zeros(self.npnt, dtype=bool)
        m[self.bf_pnt.ravel()] = True
        return m

    def boundary_edge_mask(self) -> np.ndarray:
        m = np.zeros(self.nseg, dtype=bool)
        m[self.bf_seg.ravel()] = True
        return m

    def volumes(self) -> np.ndarray:
        """tet_mesh::volume (tet_prims.cpp:180-191): (v3-v2) x (v1-v2) . (v0-v2)."""
        p = self.pnts[self.tet_pnt]
        return np.einsum("ij,ij->i", np.cross(p[:, 3] - p[:, 2], p[:, 1] - p[:, 2]), p[:, 0] - p[:, 2])


def _triple(p):
    return np.einsum("ij,ij->i", np.cross(p[:, 3] - p[:, 2], p[:, 1] - p[:, 2]), p[:, 0] - p[:, 2])


def from_tets(pnts: np.ndarray, tets: np.ndarray) -> TetMesh:
    """Edges, signs and boundary faces of a tetrahedralisation (any vertex order)."""
    pnts = np.ascontiguousarray(pnts, dtype=np.float64)
    tets = np.array(tets, dtype=np.int64)
    neg = _triple(pnts[tets]) < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    if np.any(_triple(pnts[tets]) <= 0):
        raise ValueError("degenerate tetrahedron")
    ends = tets[:, VS]                       # (ntet, 6, 2)
    lo, hi = ends.min(axis=2), ends.max(axis=2)
    keys = lo * pnts.shape[0] + hi
    uk, inv = np.unique(keys.ravel(), return_inverse=True)
    seg_pnt = np.stack([uk // pnts.shape[0], uk % pnts.shape[0]], axis=1)
    tet_seg = inv.reshape(-1, 6)
    tet_sgn = np.where(ends[:, :, 0] == seg_pnt[tet_seg][:, :, 0], 1, -1)
    # boundary faces: faces of exactly one tet, ordered as that tet orders them
    faces = tets[:, VF]                      # (ntet, 4, 3)
    fs = np.sort(faces, axis=2)
    n = pnts.shape[0]
    fk = (fs[:, :, 0] * n + fs[:, :, 1]) * n + fs[:, :, 2]
    u, first, cnt = np.unique(fk.ravel(), return_index=True, return_counts=True)
    bsel = first[cnt == 1]
    bf_pnt = faces.reshape(-1, 3)[bsel]
    seg_index = {int(k): i for i, k in enumerate(uk)}
    bf_seg = np.zeros_like(bf_pnt)
    bf_sgn = np.zeros_like(bf_pnt)
    for j in range(3):
        a, b = bf_pnt[:, (j + 1) % 3], bf_pnt[:, (j + 2) % 3]
        ks = np.minimum(a, b) * n + np.maximum(a, b)
        bf_seg[:, j] = [seg_index[int(k)] for k in ks]
        bf_sgn[:, j] = np.where(seg_pnt[bf_seg[:, j], 0] == a, 1, -1)
    return TetMesh(pnts, seg_pnt.astype(np.int32), tets.astype(np.int32), tet_seg.astype(np.int32),
                   tet_sgn.astype(np.int32), bf_pnt.astype(np.int32), bf_seg.astype(np.int32),
                   bf_sgn.astype(np.int32))


def unit_cube(n: int, jitter: float = 0.0, seed: int = 0) -> TetMesh:
    """n^3 cubes, each split into 6 tetrahedra along its main diagonal (Kuhn);
    interior vertices jittered U(-jitter h, jitter h) per coordinate."""
    N1 = n + 1
    h = 1.0 / n
    g = np.arange(N1) * h
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    pnts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    if jitter:
        rng = np.random.default_rng(seed)
        ii = np.arange(N1)
        kk, jj, i0 = np.meshgrid(ii, ii, ii, indexing="ij")
        interior = ((i0 > 0) & (i0 < n) & (jj > 0) & (jj < n) & (kk > 0) & (kk < n)).ravel()
        pnts[interior] += rng.uniform(-jitter * h, jitter * h, size=(interior.sum(), 3))

    def vid(i, j, k):
        return i + N1 * (j + N1 * k)

    ci, cj, ck = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    tets = []
    for perm in permutations(range(3)):
        steps = [np.zeros(3, dtype=int)]
        for a in perm:
            s = steps[-1].copy()
            s[a] = 1
            steps.append(s)
        tets.append(np.stack([vid(ci + s[0], cj + s[1], ck + s[2]) for s in steps], axis=1))
    tets = np.stack(tets, axis=1).reshape(-1, 4)
    return from_tets(pnts, tets)


def check_mesh(m: TetMesh) -> None:
    """Invariants the reference relies on: positive orientation, consistent edge
    signs, outward boundary-face ordering, closed boundary.  Raises ValueError naming
    the first one that fails."""
    def require(ok, what):
        if not bool(np.all(ok)):
            raise ValueError(f"inconsistent tet mesh: {what}")

    require(m.volumes() > 0, "tetrahedra must be positively oriented")
    ends = m.tet_pnt[:, VS]
    sp = m.seg_pnt[m.tet_seg]
    fwd = (ends[:, :, 0] == sp[:, :, 0]) & (ends[:, :, 1] == sp[:, :, 1])
    bwd = (ends[:, :, 0] == sp[:, :, 1]) & (ends[:, :, 1] == sp[:, :, 0])
    require(fwd | bwd, "an edge of a tet does not join its vertices")
    require(np.where(fwd, 1, -1) == m.tet_sgn, "tet_sgn disagrees with seg_pnt")
    p = m.pnts[m.bf_pnt]
    nrm = np.cross(p[:, 2] - p[:, 1], p[:, 0] - p[:, 1])
    c = p.mean(axis=1) - m.pnts.mean(axis=0)
    require(np.einsum("ij,ij->i", nrm, c) > 0, "a boundary face is not ordered outward")
    require(np.allclose(nrm.sum(axis=0), 0.0, atol=1e-12), "the boundary is not closed")
