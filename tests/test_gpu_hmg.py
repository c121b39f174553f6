"""GPU h-multigrid (tri_hp/movco.cpp:41-102, tri_hp/movfin.cpp:16-82, block::cycle in
tri_mesh/blocks.cpp:1267-1348) against the oracle, through the C-ABI.

Transfers to 1e-13 relative; whole h+p cycles to 1e-11 relative after several cycles;
the h-levels must cut the cycles to 1e-10 by more than half and reach the solution
p-multigrid alone reaches (modulo the constant-pressure null mode)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
INFLOW = 0b1110


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _stacks(oracle_mod, physics, n, log2p, ngrid):
    from hpmp import mesh as M, multigrid as MG, workloads as W
    from hpmp.tri_hp import TriHpCns, TriHpIns
    p = 1 << log2p
    if physics == "ins":
        m = M.unit_square(n, 0.2, 15)
        st = W.noisy_freestream(m, p, seed=16, scale=0.3)
        prm = dict(mu=5e-2)
        gcls, ocls = TriHpIns, oracle_mod.TriHpIns
    else:
        m = M.unit_square(n, 0.2, 3, essential_mask=INFLOW)
        st, bc = W.cns_gaussian_state(m, p, seed=11, amp=0.05, width=0.2, noise=2e-2)
        prm = dict(mu=1e-2, bc=bc, cfl=(0.01, 0.01, 0.01))
        gcls, ocls = TriHpCns, oracle_mod.TriHpCns
    meshes, links = MG.square_hierarchy(m, n, ngrid)
    G = gcls(m, log2p=log2p, **prm)
    O = ocls(m, log2p=log2p, **prm)
    G.set_state(*st)
    O.set_state(*st)
    return meshes, MG.build_stack(gcls, G, meshes, links, **prm), MG.build_stack(ocls, O, meshes, links, **prm)


@pytest.mark.parametrize("physics", ["ins", "cns"])
def test_h_transfer_parity(oracle_mod, physics):
    _case(f"h restrict/prolongate {physics} 8x8 -> 4x4 p=2")
    meshes, (G, Gc), (O, Oc) = _stacks(oracle_mod, physics, 8, 1, 2)
    rng = np.random.default_rng(5)
    res = [rng.standard_normal(a.shape) for a in O.get_res()]
    G.set_residual(*res)
    O.set_res(*res)
    Gc.mg_restrict()
    Oc.mg_restrict()
    err = rel_err(Gc.get_state(), Oc.get_state())
    print("coarse state rel err", err, flush=True)
    assert err < 1e-14
    # the first coarse residual is fadd x the restricted residual plus nothing else
    Gc.rsdl()
    Oc.rsdl()
    err = rel_err(Gc.get_residual(), Oc.get_res())
    print("coarse first residual rel err", err, flush=True)
    assert err < 1e-13
    cs = Oc.get_state()
    cm = meshes[1]
    d = np.stack([0.01 * (k + 1) * (cm.pnts[:, 0] - 0.5 * cm.pnts[:, 1]) for k in range(cs[0].shape[1])], 1)
    Gc.set_state(cs[0] + d, cs[1], cs[2])
    Oc.set_state(cs[0] + d, cs[1], cs[2])
    Gc.mg_prolongate()
    Oc.mg_prolongate()
    err = rel_err(G.get_state(), O.get_state())
    err_r = rel_err(G.get_residual()[:1], O.get_res()[:1])
    print("fine state rel err", err, "fine correction rel err", err_r, flush=True)
    assert err < 1e-14 and err_r < 1e-13


@pytest.mark.parametrize("physics,n,log2p,ngrid", [("ins", 16, 2, 3), ("ins", 8, 0, 3), ("cns", 8, 1, 2)])
def test_h_cycle_parity(oracle_mod, physics, n, log2p, ngrid):
    _case(f"h+p cycles {physics} {n}x{n} log2p={log2p} ngrid={ngrid}")
    from hpmp import multigrid as MG
    _, gg, og = _stacks(oracle_mod, physics, n, log2p, ngrid)
    for k in range(4):
        eg = MG.cycle(gg, log2p)
        eo = MG.cycle(og, log2p)
        print("cycle", k, "maxres gpu", eg, "oracle", eo, flush=True)
        assert abs(eg - eo) <= 1e-10 * abs(eo)
    err = rel_err(gg[0].get_state(), og[0].get_state())
    print("state rel err", err, flush=True)
    assert err < 1e-11


def test_h_multigrid_convergence_gain():
    _case("ins 32x32 p=2: cycles to 1e-10 with 4 grids vs 1")
    from hpmp import mesh as M, multigrid as MG, workloads as W
    from hpmp.tri_hp import TriHpIns, rel_diff_mod_pressure
    n, log2p = 32, 1
    m = M.unit_square(n, 0.2, 15)
    st = W.noisy_freestream(m, 2, seed=16, scale=0.3)
    meshes, links = MG.square_hierarchy(m, n, 4)
    out = {}
    for ng in (4, 1):
        F = TriHpIns(m, log2p=log2p, mu=5e-2)
        F.set_state(*st)
        grids = MG.build_stack(TriHpIns, F, meshes[:ng], links[:ng - 1], mu=5e-2)
        e0 = None
        for it in range(8000):
            e = MG.cycle(grids, log2p)
            e0 = e if e0 is None else e0
            if e < 1e-10 * e0:
                break
        out[ng] = (it + 1, F.get_state(), e / e0)
        print("ngrid", ng, "cycles", it + 1, "rel maxres", e / e0, flush=True)
    assert out[4][2] < 1e-10 and out[1][2] < 1e-10
    assert 2 * out[4][0] < out[1][0]
    err = rel_diff_mod_pressure(out[4][1], out[1][1])
    print("solution rel diff (mod pressure)", err, flush=True)
    assert err < 1e-8


@pytest.mark.parametrize("physics,n,log2p,ngrid,vw,ic,ir", [("ins", 16, 2, 3, 2, 2, 2), ("cns", 8, 1, 2, 1, 1, 1)])
def test_fused_cycle_is_the_call_by_call_cycle(oracle_mod, physics, n, log2p, ngrid, vw, ic, ir):
    """hpg_mg_cycle enqueues the same kernels in the same order as the level-by-level
    recursion: states and maxres bitwise equal; and it stays on the oracle's trajectory."""
    _case(f"fused cycle vs call by call {physics} {n}x{n} log2p={log2p} ngrid={ngrid} vw={vw} {ic}/{ir}")
    from hpmp import multigrid as MG
    _, ga, og = _stacks(oracle_mod, physics, n, log2p, ngrid)
    _, gb, _ = _stacks(oracle_mod, physics, n, log2p, ngrid)
    for k in range(3):
        ea = MG.cycle(ga, log2p, vw, ic, ir, fused=True)
        eb = MG.cycle(gb, log2p, vw, ic, ir, fused=False)
        eo = MG.cycle(og, log2p, vw, ic, ir)
        print("cycle", k, "maxres fused", ea, "call by call", eb, "oracle", eo, flush=True)
        assert ea == eb
        assert abs(ea - eo) <= 1e-10 * abs(eo)
    for x, y in zip(ga[0].get_state(), gb[0].get_state()):
        assert np.array_equal(x, y)
    err = rel_err(ga[0].get_state(), og[0].get_state())
    print("state rel err vs oracle", err, flush=True)
    assert err < 1e-11


@pytest.mark.parametrize("physics,form", [("ins", "fp64"), ("ins", "mixed"), ("ins", "learned"), ("cns", "fp64")])
def test_cycle_graph_replays_the_enqueued_cycle(oracle_mod, physics, form):
    """hpg_mg_cycle captures the cycle into a HIP graph at the second call with unchanged
    settings and replays it afterwards: every replay must give the state and maxres of
    the enqueued cycle bitwise, with eager work (a residual evaluation and a host read)
    between the replays; a changed setting (the CFL) must not replay the old graph."""
    _case(f"cycle graph replay vs enqueued cycle {physics} {form}")
    from hpmp import _lib, multigrid as MG
    n, log2p, ngrid = (16, 2, 3) if physics == "ins" else (8, 1, 2)
    vw, ic, ir = (2, 2, 2) if physics == "ins" else (1, 1, 1)
    _, ga, _ = _stacks(oracle_mod, physics, n, log2p, ngrid)
    _, gb, _ = _stacks(oracle_mod, physics, n, log2p, ngrid)
    ga[0].set_cycle_graph(False)
    for g in (ga, gb):
        if form == "mixed":
            g[0].set_precision(_lib.HPG_FP32_DEFECT, ninner=1)
        if form == "learned":
            for x in g:
                m = x.mesh
                y = np.tile(np.log([0.5, 0.5, 0.5, 0.7, 0.7, 0.7, 1.0, 1.0, 1.0]), (m.ntri, 1)).astype(np.float32)
                assert x.L.hpg_set_prcn_log_scalings(x.h, _lib.ptr(y)) == 0
                x.set_preconditioner_mode(_lib.HPG_PRCN_LEARNED)
    for k in range(6):
        if k == 4:
            # a new CFL list: the next call runs enqueued, the one after captures anew
            new_cfl = [0.8 * c for c in (2.5, 1.5, 1.0, 0.5)] if physics == "ins" else [0.008] * 4
            for g in (ga, gb):
                for x in g:
                    x.set_cfl(new_cfl)
        ea = MG.cycle(ga, log2p, vw, ic, ir, fused=True)
        eb = MG.cycle(gb, log2p, vw, ic, ir, fused=True)
        ra = ga[0].get_state()
        rb = gb[0].get_state()
        ga[0].rsdl()
        gb[0].rsdl()
        print("cycle", k, "maxres enqueued", ea, "graph", eb, "judge", ga[0].resnorm(), gb[0].resnorm(), flush=True)
        assert ea == eb
        for x, y in zip(ra, rb):
            assert np.array_equal(x, y)
        assert np.array_equal(ga[0].get_residual()[0], gb[0].get_residual()[0])
