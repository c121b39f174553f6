"""h-multigrid transfers of the oracle (tri_hp/movco.cpp:41-102, tri_hp/movfin.cpp:16-82)
and the mesh connections they use (tri_mesh/tri_mesh.h:197-201).

Pins: the connections reproduce vertex positions (barycentric weights); restriction is
the transpose of linear interpolation, so it preserves the residual's moments against
every linear function; prolongation of a linear coarse correction adds exactly that
linear field to the fine vertices; a two- and three-grid cycle converges to the same
discrete solution as p-multigrid alone, in fewer cycles.
"""
import numpy as np
import pytest

from hpmp import mesh as M, multigrid as MG, workloads as W
from hpmp.tri_hp import rel_diff_mod_pressure


def _lin(pts, a):
    return a[0] + a[1] * pts[:, 0] + a[2] * pts[:, 1]


@pytest.mark.parametrize("n", [8, 12])
def test_connections_reproduce_positions(n):
    m = M.unit_square(n, 0.2, 5)
    meshes, links = MG.square_hierarchy(m, n, 2)
    f, c = meshes
    cc_tri, cc_wt, fc_tri, fc_wt = links[0]
    assert c.npnt == (n // 2 + 1) ** 2
    assert np.allclose(cc_wt.sum(1), 1.0, atol=1e-15) and cc_wt.min() > -1e-14
    assert np.allclose(fc_wt.sum(1), 1.0, atol=1e-15) and fc_wt.min() > -1e-14
    xr = np.einsum("ij,ijk->ik", cc_wt, c.pnts[c.tri_pnt[cc_tri]])
    assert np.abs(xr - f.pnts).max() < 1e-14
    xr = np.einsum("ij,ijk->ik", fc_wt, f.pnts[f.tri_pnt[fc_tri]])
    assert np.abs(xr - c.pnts).max() < 1e-14


def _pair(oracle_mod, cls_name, n=8, log2p=1, seed=3):
    m = M.unit_square(n, 0.2, seed)
    meshes, links = MG.square_hierarchy(m, n, 2)
    cls = getattr(oracle_mod, cls_name)
    p = 1 << log2p
    if cls_name == "TriHpIns":
        st = W.noisy_freestream(m, p, seed=seed, scale=0.3)
        F = cls(m, log2p=log2p, mu=5e-2)
        grids = MG.build_stack(cls, F, meshes, links, mu=5e-2)
    else:
        st, bc = W.cns_freestream(m, p, mach=0.3)
        st = (st[0] * (1 + 0.01 * np.random.default_rng(seed).standard_normal(st[0].shape)), st[1], st[2])
        F = cls(m, log2p=log2p, bc=bc)
        grids = MG.build_stack(cls, F, meshes, links, bc=bc)
    F.set_state(*st)
    return grids, meshes, links


@pytest.mark.parametrize("cls_name", ["TriHpIns", "TriHpCns"])
def test_restriction_is_transpose_of_interpolation(oracle_mod, cls_name):
    (F, Cg), (fm, cm), _ = _pair(oracle_mod, cls_name)
    nv = F.get_state()[0].shape[1]
    rng = np.random.default_rng(7)
    rv = rng.standard_normal((fm.npnt, nv))
    F.set_res(rv, *F.get_res()[1:])
    Cg.mg_restrict()
    r0, vug = Cg.get_mg()
    for a in ([1.0, 0.0, 0.0], [0.3, 1.0, 0.0], [-0.2, 0.4, 1.3]):
        lhs = (r0 * _lin(cm.pnts, a)[:, None]).sum(0)
        rhs = (rv * _lin(fm.pnts, a)[:, None]).sum(0)
        assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(rv).sum()
    # coarse vertices are fine vertices: interpolated state equals the fine state there
    fv = F.get_state()[0]
    cvv = Cg.get_state()[0]
    n = round(np.sqrt(fm.npnt)) - 1
    N1c = n // 2 + 1
    jj, ii = np.divmod(np.arange(N1c * N1c), N1c)
    assert np.abs(cvv - fv[2 * jj * (n + 1) + 2 * ii]).max() < 1e-14 * np.abs(fv).max()
    assert np.array_equal(vug, cvv)


@pytest.mark.parametrize("cls_name", ["TriHpIns", "TriHpCns"])
def test_prolongation_adds_linear_correction(oracle_mod, cls_name):
    (F, Cg), (fm, cm), _ = _pair(oracle_mod, cls_name)
    fv0 = F.get_state()
    Cg.mg_restrict()
    cs = Cg.get_state()
    nv = cs[0].shape[1]
    coef = [[0.1 * k, -0.2 + 0.05 * k, 0.3 - 0.1 * k] for k in range(nv)]
    d_c = np.stack([_lin(cm.pnts, c) for c in coef], 1)
    Cg.set_state(cs[0] + d_c, cs[1], cs[2])
    Cg.mg_prolongate()
    d_f = np.stack([_lin(fm.pnts, c) for c in coef], 1)
    fv1 = F.get_state()
    assert np.abs(fv1[0] - fv0[0] - d_f).max() < 1e-14
    assert np.abs(F.get_res()[0] - d_f).max() < 1e-14
    # only vertex values move on the fine mesh
    assert np.array_equal(fv1[1], fv0[1]) and np.array_equal(fv1[2], fv0[2])


def test_h_cycle_converges_to_the_same_solution_faster(oracle_mod):
    n, log2p, tol = 8, 1, 1e-9
    m = M.unit_square(n, 0.2, 15)
    st = W.noisy_freestream(m, 2, seed=16, scale=0.3)
    meshes, links = MG.square_hierarchy(m, n, 3)
    out = {}
    for ng in (1, 3):
        F = oracle_mod.TriHpIns(m, log2p=log2p, mu=5e-2)
        F.set_state(*st)
        grids = MG.build_stack(oracle_mod.TriHpIns, F, meshes[:ng], links[:ng - 1], mu=5e-2)
        e0 = None
        for it in range(3000):
            e = MG.cycle(grids, log2p)
            e0 = e if e0 is None else e0
            if e < tol * e0:
                break
        out[ng] = (it + 1, F.get_state())
    print("cycles p-only", out[1][0], "h+p", out[3][0])
    assert out[3][0] * 2 < out[1][0]
    assert rel_diff_mod_pressure(out[1][1], out[3][1]) < 1e-8
