"""Periodic boundaries (the reference's prdc pairs) in the oracle.

The oracle restates the reference's treatment: every copy of a periodic vertex or side
keeps its partial residual, the preconditioner sums and the solved values are averaged
over the copies where tri_hp exchanges them (tri_hp/setup_preconditioner.cpp:24-36,
171-176, 255-258, 331-334).  What fixes the result independently of that restatement:
a strip periodic in y is the same problem as the strip repeated twice, so from a state
with period 1 the relaxation on the single strip and on the doubled strip must produce
the same state at corresponding points (the doubled strip assembles the periodic
seam as ordinary interior vertices and sides).
"""
import numpy as np
import pytest


def periodic_strip(n, copies=1, jitter=0.0, seed=0, masks=(0b000, 0b000, 0b000, 0b011)):
    """[0, 1] x [0, copies] with n x (n * copies) squares cut along alternating
    diagonals; edge boundaries bottom (0), right (1), top (2), left (3) with `masks`;
    bottom and top pair periodically.  Interior vertices are jittered with one pattern
    repeated in every copy, so the strip is periodic with period 1."""
    from hpmp.coarsen import build_mesh
    rng = np.random.default_rng(seed)
    N1 = n + 1
    J = rng.uniform(-0.5, 0.5, (N1, n, 2)) * jitter / n   # one period of jitter (rows 0..n-1)
    rows = n * copies
    ii, jj = np.meshgrid(np.arange(N1), np.arange(rows + 1), indexing="ij")
    X = ii / n + 0.0
    Y = jj / n + 0.0
    jx = J[ii, jj % n, 0]
    jy = J[ii, jj % n, 1]
    inner_x = (ii > 0) & (ii < n)
    X = X + np.where(inner_x, jx, 0.0)
    Y = Y + np.where(inner_x & (jj % n != 0), jy, 0.0)   # the rows y = 0, 1, ... stay straight
    pnts = np.stack([X.ravel(), Y.ravel()], axis=1)

    def vid(i, j):
        return i * (rows + 1) + j

    tri = []
    for i in range(n):
        for j in range(rows):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            tri += [(a, b, c), (a, c, d)] if (i + j % n) % 2 == 0 else [(a, b, d), (b, c, d)]
    eps = 1e-9

    def side_bdry(u, v):
        (x0, y0), (x1, y1) = pnts[u], pnts[v]
        if abs(y0) < eps and abs(y1) < eps:
            return 0
        if abs(x0 - 1) < eps and abs(x1 - 1) < eps:
            return 1
        if abs(y0 - rows / n) < eps and abs(y1 - rows / n) < eps:
            return 2
        return 3

    return build_mesh(pnts, np.array(tri), side_bdry, 4, list(masks), names=["bottom", "right", "top", "left"],
                      ids=[1, 2, 1, 4])


def half_map(single, double, p):
    """For every vertex, side and element of `single`, its counterpart in the lower copy
    of `double` (the same points; the periodic top row of `single` maps to y = 1 of
    `double`), and the relative direction of the sides."""
    from scipy.spatial import cKDTree
    tv = cKDTree(double.pnts)
    _, vmap = tv.query(single.pnts)
    mid_s = 0.5 * (single.pnts[single.seg_pnt[:, 0]] + single.pnts[single.seg_pnt[:, 1]])
    mid_d = 0.5 * (double.pnts[double.seg_pnt[:, 0]] + double.pnts[double.seg_pnt[:, 1]])
    _, smap = cKDTree(mid_d).query(mid_s)
    same = vmap[single.seg_pnt[:, 0]] == double.seg_pnt[smap, 0]
    cen_s = single.pnts[single.tri_pnt].mean(axis=1)
    cen_d = double.pnts[double.tri_pnt].mean(axis=1)
    _, tmap = cKDTree(cen_d).query(cen_s)
    return vmap, smap, np.where(same, 1.0, -1.0), tmap


def periodic_state(m, p, seed, amp=0.2):
    """A state of period 1 in y (smooth in x, y): vertex and side values from a Fourier
    field in the vertices' positions; side and interior modes small random numbers
    replicated by position (so copies agree)."""
    sm, im = p - 1, (p - 1) * (p - 2) // 2
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((3, 4))

    def field(x):
        y = 2 * np.pi * x[:, 1]
        return np.stack([1.0 + amp * (c[n, 0] * np.sin(y) + c[n, 1] * np.cos(y) * x[:, 0] + c[n, 2] * np.sin(2 * y)
                                      + c[n, 3] * x[:, 0] ** 2) for n in range(3)], axis=1)

    v = field(m.pnts)
    v[:, 1] -= 1.0
    mid = 0.5 * (m.pnts[m.seg_pnt[:, 0]] + m.pnts[m.seg_pnt[:, 1]])
    s = np.zeros((m.nseg, sm, 3))
    for k in range(sm):
        s[:, k, :] = 0.05 * (field(mid + 0.01 * (k + 1)) - 1.0)
    cen = m.pnts[m.tri_pnt].mean(axis=1)
    i = np.zeros((m.ntri, im, 3))
    for k in range(im):
        i[:, k, :] = 0.02 * (field(cen + 0.013 * (k + 1)) - 1.0)
    return v, s, i


def to_double(st, single, double, p, maps):
    """The single strip's state copied to both copies of the doubled strip (by position)."""
    vmap, smap, sdir, tmap = maps
    sm, im = p - 1, (p - 1) * (p - 2) // 2
    v = np.zeros((double.npnt, 3)); s = np.zeros((double.nseg, sm, 3)); i = np.zeros((double.ntri, im, 3))
    sign = np.array([(-1.0) ** k for k in range(sm)])
    for shift in (0.0, 1.0):
        pts = single.pnts + (0.0, shift)
        from scipy.spatial import cKDTree
        _, vm = cKDTree(double.pnts).query(pts)
        v[vm] = st[0]
        mid_s = 0.5 * (pts[single.seg_pnt[:, 0]] + pts[single.seg_pnt[:, 1]])
        mid_d = 0.5 * (double.pnts[double.seg_pnt[:, 0]] + double.pnts[double.seg_pnt[:, 1]])
        _, sm_ = cKDTree(mid_d).query(mid_s)
        same = vm[single.seg_pnt[:, 0]] == double.seg_pnt[sm_, 0]
        s[sm_] = np.where(same[:, None, None], st[1], st[1] * sign[None, :, None])
        cen_s = pts[single.tri_pnt].mean(axis=1)
        _, tm = cKDTree(double.pnts[double.tri_pnt].mean(axis=1)).query(cen_s)
        i[tm] = st[2]
    return v, s, i


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_periodic_strip_equals_the_doubled_strip(oracle_mod, log2p):
    n, p = 4, 1 << log2p
    A = periodic_strip(n, 1, jitter=0.3, seed=3)
    B = periodic_strip(n, 2, jitter=0.3, seed=3)
    st = periodic_state(A, p, seed=5)
    maps = half_map(A, B, p)
    OA = oracle_mod.TriHpIns(A, log2p=log2p, mu=2e-2)
    OB = oracle_mod.TriHpIns(B, log2p=log2p, mu=2e-2)
    OA.set_periodic([(0, 2)])
    OB.set_periodic([(0, 2)])
    OA.set_state(*st)
    OB.set_state(*to_double(st, A, B, p, maps))
    for _ in range(3):
        assert OA.setup_preconditioner() == 0 and OB.setup_preconditioner() == 0
        OA.update()
        OB.update()
    a, b = OA.get_state(), OB.get_state()
    vmap, smap, sdir, tmap = maps
    sgn = np.array([[1.0 if (k % 2 == 0 or d > 0) else -1.0 for k in range(p - 1)] for d in sdir])
    dv = np.abs(a[0] - b[0][vmap]).max()
    ds = np.abs(a[1] - b[1][smap] * sgn[:, :, None]).max() if p > 1 else 0.0
    di = np.abs(a[2] - b[2][tmap]).max() if a[2].size else 0.0
    scale = np.abs(a[0]).max()
    print("single vs doubled strip after 3 updates:", dv, ds, di)
    assert max(dv, ds, di) < 1e-13 * scale
    # the state moved (the test is not vacuous)
    assert np.abs(a[0] - st[0]).max() > 1e-4


def test_periodic_pairing_rejects_mismatched_boundaries(oracle_mod):
    A = periodic_strip(4, 1)
    O = oracle_mod.TriHpIns(A, log2p=1, mu=1e-2)
    with pytest.raises(RuntimeError):
        O.set_periodic([(0, 1)])   # bottom and right: not one translation apart
