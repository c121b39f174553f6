"""Curved boundary sides: the coordinate side modes of a boundary that follows a shape,
and the side modes of boundary data along such a side.

`side_modes` restates hp_edge_bdry::curv_init (tri_hp/hp_boundary.cpp:621-675): for
every side it places the side's Gauss points on the straight chord between its end
vertices (proj1d of the vertex coordinates), moves each onto the boundary
(`to_boundary`, the role of edge_bdry::mvpttobdry), integrates the chord-minus-boundary
offset against the side modes (intgrt1d) and solves the side-mode mass system (the
reference's banded dpbtrs with sdiag1d): the side modes are minus that solution, so
that vertex modes + side modes reproduce the boundary's L2 projection.  The result is
the ``crds`` array that TriHpIns.set_curved (hpg_set_curved) takes.

`side_data_modes` projects boundary data the same way (the reference's essential
boundary values on a side: the vertex values and the side modes of the data's
deviation from their linear interpolant), at the points of the curved side.

`circle` is a `to_boundary` for a circular boundary; `annulus` builds a triangulated
ring between two circles with the boundary ids of its inner and outer circles.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .mesh import TriMesh


def _edge_tables(p: int):
    """(g1, w1): the 1-D basis at the p + 1 Gauss points, gpx x (p + 1) with columns
    (vertex 0, vertex 1, side modes), and the Gauss weights."""
    L = _lib.load()
    n = L.hpg_basis_operator(p, b"g1", None)
    g1 = np.zeros(n)
    L.hpg_basis_operator(p, b"g1", g1.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)))
    w = np.zeros(L.hpg_basis_operator(p, b"wx", None))
    L.hpg_basis_operator(p, b"wx", w.ctypes.data_as(_lib.C.POINTER(_lib.C.c_double)))
    return g1.reshape(p + 1, p + 1), w


def _project_offsets(g1, w, off):
    """Side modes c with sum_m c_m phi_m = -(L2 projection of `off` onto the side modes);
    off: (nside, gpx, k) values at the Gauss points."""
    gs = g1[:, 2:]
    M = (gs * w[:, None]).T @ gs
    rhs = np.einsum("i,im,sik->smk", w, gs, off)
    return -np.linalg.solve(M[None], rhs) if gs.shape[1] else np.zeros(off.shape[:1] + (0,) + off.shape[2:])


def side_modes(mesh: TriMesh, sides, to_boundary, p: int) -> np.ndarray:
    """crds (nside, p - 1, 2) of the listed boundary sides for the degree-p basis;
    ``to_boundary(x)`` maps points (n, 2) onto the boundary (the side's end vertices must
    already lie on it)."""
    sides = np.asarray(sides, dtype=np.int64)
    if p < 2 or sides.size == 0:
        return np.zeros((sides.size, max(p - 1, 0), 2))
    g1, w = _edge_tables(p)
    X = mesh.pnts[mesh.seg_pnt[sides]]                  # (nside, 2 vertices, 2)
    chord = np.einsum("iv,svk->sik", g1[:, :2], X)      # straight-side Gauss points
    on = np.asarray(to_boundary(chord.reshape(-1, 2)), dtype=np.float64).reshape(chord.shape)
    return _project_offsets(g1, w, chord - on)


def side_points(mesh: TriMesh, sides, crds, p: int) -> np.ndarray:
    """Positions (nside, p + 1, 2) of the 1-D Gauss points on the listed sides with their
    coordinate side modes (zeros: straight)."""
    g1, _ = _edge_tables(p)
    sides = np.asarray(sides, dtype=np.int64)
    X = mesh.pnts[mesh.seg_pnt[sides]]
    pts = np.einsum("iv,svk->sik", g1[:, :2], X)
    if p >= 2:
        pts = pts + np.einsum("im,smk->sik", g1[:, 2:], np.asarray(crds).reshape(sides.size, p - 1, 2))
    return pts


def side_data_modes(mesh: TriMesh, sides, crds, f, p: int) -> np.ndarray:
    """Side modes (nside, p - 1, k) of the boundary data ``f(x) -> (n, k)`` along the listed
    (curved) sides: the vertex values are f at the end vertices; the side modes carry the
    L2 projection of f's deviation from their linear interpolant."""
    sides = np.asarray(sides, dtype=np.int64)
    g1, w = _edge_tables(p)
    pts = side_points(mesh, sides, crds, p)
    fv = np.asarray(f(pts.reshape(-1, 2)), dtype=np.float64)
    k = fv.shape[-1] if fv.ndim > 1 else 1
    fv = fv.reshape(sides.size, p + 1, k)
    ends = np.asarray(f(mesh.pnts[mesh.seg_pnt[sides]].reshape(-1, 2)), dtype=np.float64).reshape(sides.size, 2, k)
    lin = np.einsum("iv,svk->sik", g1[:, :2], ends)
    return _project_offsets(g1, w, lin - fv)


def circle(cx: float, cy: float, r: float):
    """`to_boundary` of the circle of radius r about (cx, cy): radial projection."""
    def to_boundary(x):
        d = np.asarray(x, dtype=np.float64) - (cx, cy)
        return (cx, cy) + r * d / np.linalg.norm(d, axis=1, keepdims=True)
    return to_boundary


def annulus(nr: int, nt: int, r0: float = 0.5, r1: float = 1.0, inner_mask: int = 0b011, outer_mask: int = 0b011,
            jitter: float = 0.0, seed: int = 0) -> TriMesh:
    """Triangulated ring r0 <= r <= r1 about the origin: nr cells across, nt around, every
    quadrilateral cut along a diagonal that alternates with position.  Edge boundary 0 is
    the inner circle (id 1), 1 the outer circle (id 2), with the given essential masks;
    both run counter-clockwise around the domain.  Interior vertices can be jittered by
    ``jitter`` times the local cell size."""
    from .coarsen import build_mesh
    rng = np.random.default_rng(seed)
    rr = np.linspace(r0, r1, nr + 1)
    th = np.linspace(0.0, 2 * np.pi, nt, endpoint=False)
    R, T = np.meshgrid(rr, th, indexing="ij")
    if jitter > 0.0:
        dr = (r1 - r0) / nr
        R[1:-1] += jitter * dr * rng.uniform(-0.5, 0.5, R[1:-1].shape)
        T[1:-1] += jitter * (2 * np.pi / nt) * rng.uniform(-0.5, 0.5, T[1:-1].shape)
    pnts = np.stack([R * np.cos(T), R * np.sin(T)], axis=-1).reshape(-1, 2)
    vid = np.arange((nr + 1) * nt).reshape(nr + 1, nt)
    tri = []
    for i in range(nr):
        for j in range(nt):
            a, b = vid[i, j], vid[i, (j + 1) % nt]
            c, d = vid[i + 1, j], vid[i + 1, (j + 1) % nt]
            if (i + j) % 2 == 0:
                tri += [(a, b, d), (a, d, c)]
            else:
                tri += [(a, b, c), (b, d, c)]
    inner = set(vid[0].tolist())

    def side_bdry(u, v):
        return 0 if (u in inner and v in inner) else 1

    return build_mesh(pnts, np.array(tri), side_bdry, 2, [inner_mask, outer_mask], names=["inner", "outer"], ids=[1, 2])
