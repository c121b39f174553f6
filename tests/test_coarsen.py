"""Coarse meshes of unstructured meshes for h-multigrid (hpmp.coarsen) and their
connection tables, on the meshes of the reference's own test decks.

The tables are pinned by what they must reproduce: every coarse vertex is a fine
vertex, so its fine-triangle weights must pick it exactly; every fine vertex inside
the coarse mesh must be the weighted sum of its coarse triangle's corners; weights
are non-negative and sum to one.  The oracle then runs the h+p cycle on the
hierarchy: it must converge, and in fewer cycles than p-multigrid alone."""
import os

import numpy as np
import pytest

from hpmp import grd, mesh as M
from hpmp.coarsen import coarsen, connect, hierarchy, locate

HERE = os.path.join(os.path.dirname(__file__), "golden", "grd")


@pytest.mark.parametrize("name", ["tim.1", "channel", "square4"])
def test_coarse_meshes_and_tables(name):
    m = grd.read_grd(os.path.join(HERE, name + ".grd"), {1: 0b011, 2: 0b001, 4: 0b011, 8: 0b010, 640: 0b011})
    meshes, links = hierarchy(m, 4)
    assert len(meshes) >= 2
    for a, b, (cc_tri, cc_wt, fc_tri, fc_wt) in zip(meshes[:-1], meshes[1:], links):
        assert b.ntri < a.ntri and b.npnt < a.npnt
        # boundaries keep their ids and essential masks
        assert b.nebd == a.nebd
        assert np.array_equal(b.ebdry_essential, a.ebdry_essential)
        assert b.ebdry_ids == a.ebdry_ids
        # every boundary vertex of the coarse mesh lies on the fine boundary
        fb = a.pnts[a.boundary_vertex_mask()]
        for p in b.pnts[b.boundary_vertex_mask()]:
            assert np.min(np.abs(fb - p).sum(axis=1)) < 1e-14
        for w in (cc_wt, fc_wt):
            assert np.all(w >= -1e-12) and np.allclose(w.sum(axis=1), 1.0, atol=1e-13)
        rec = (a.pnts[a.tri_pnt[fc_tri]] * fc_wt[:, :, None]).sum(axis=1)
        assert np.abs(rec - b.pnts).max() < 1e-13
        _, _, inside = locate(b, a.pnts)
        rec = (b.pnts[b.tri_pnt[cc_tri]] * cc_wt[:, :, None]).sum(axis=1)
        assert np.abs(rec - a.pnts)[inside].max() < 1e-12
        # the coarse mesh covers most of the fine domain
        assert inside.mean() > 0.9


def test_jittered_square_hierarchy():
    m = M.unit_square(32, 0.2, 3)
    meshes, links = hierarchy(m, 5)
    assert [x.ntri for x in meshes][:2] == [2048, meshes[1].ntri]
    assert all(b.ntri < a.ntri for a, b in zip(meshes, meshes[1:]))
    # coarse area equals the domain's (convex domain: Delaunay covers it)
    for x in meshes:
        assert abs(0.5 * x.area2().sum() - 1.0) < 1e-12


def test_h_cycle_on_an_unstructured_hierarchy_contracts_faster(oracle_mod):
    """The oracle's h+p cycle (block::cycle, tri_mesh/blocks.cpp:1267-1348) on tim.1
    with the coarse meshes of `hierarchy` reduces the FP64 residual over 40 cycles by
    more than p-multigrid alone does (both from the same state)."""
    from hpmp import multigrid as MG, workloads as W
    m = grd.read_grd(os.path.join(HERE, "tim.1.grd"))
    st = W.noisy_freestream(m, 2, seed=3, scale=0.3)
    meshes, links = hierarchy(m, 3)
    O = oracle_mod.TriHpIns(m, log2p=1, mu=5e-2)
    grids = MG.build_stack(oracle_mod.TriHpIns, O, meshes, links, mu=5e-2)
    bv, bs = m.boundary_vertex_mask(), m.boundary_segment_mask()

    def judge(X):
        X.rsdl()
        v, s, i = [np.array(a) for a in X.get_res()]
        v[bv, :2] = 0.0
        s[bs, :, :2] = 0.0
        return max(np.abs(v).max(), np.abs(s).max())

    def reduction(gr, cycles=40):
        gr[0].set_state(*st)
        r0 = judge(gr[0])
        for _ in range(cycles):
            MG.cycle(gr, 1, vw=1, itercrsn=1, iterrfne=1)
        return judge(gr[0]) / r0

    rh = reduction(grids)
    rp = reduction([O])
    print("residual reduction over 40 cycles: h+p", rh, "p only", rp, flush=True)
    assert np.isfinite(rh) and rh < 0.5 * rp
