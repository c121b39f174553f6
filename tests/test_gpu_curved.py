"""Curved elements on the GPU (hpg_set_curved): the curved branches of element_rsdl
(tri_hp/ins/rsdl.cpp:83-257), setup_preconditioner (tri_hp/ins/setup_preconditioner.cpp:
44-90) and the boundary-edge flux (tri_hp/hp_boundary.cpp:973-1015) against the oracle
on a ring whose two circles are curved boundaries, and a converged Taylor-Couette flow
whose error against the exact solution the curved sides cut by orders of magnitude.
"""
import numpy as np
import pytest

from conftest import rel_err
from test_oracle_curved import _onb, couette_problem, linear_field_state

pytestmark = pytest.mark.gpu
RSDL_TOL = 1e-12


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _ring(oracle_mod, log2p, nr=4, nt=16, jitter=0.3, seed=2, mu=1e-2, curved=True):
    from hpmp import curved as CV
    from hpmp.tri_hp import TriHpIns
    m = CV.annulus(nr, nt, jitter=jitter, seed=seed)
    p = 1 << log2p
    G = TriHpIns(m, log2p=log2p, mu=mu)
    O = oracle_mod.TriHpIns(m, log2p=log2p, mu=mu)
    crds = CV.side_modes(m, m.ebdry_seg, _onb(), p)
    if curved:
        G.set_curved(m.ebdry_seg, crds)
        O.set_curved(m.ebdry_seg, crds)
    return m, crds, G, O


@pytest.mark.parametrize("log2p", [1, 2])
@pytest.mark.parametrize("stage", [5, 0, 1, 2])
def test_curved_rsdl_parity(oracle_mod, log2p, stage):
    _case(f"curved rsdl log2p={log2p} stage={stage}")
    from hpmp import workloads as W
    m, crds, G, O = _ring(oracle_mod, log2p)
    st = W.noisy_freestream(m, 1 << log2p, seed=11, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    for s in range(stage + 1) if stage < 5 else [5]:
        G.rsdl(s); O.rsdl(s)
    assert rel_err(G.get_residual(), O.get_res()) < RSDL_TOL


@pytest.mark.parametrize("log2p", [1, 2])
def test_curved_setup_preconditioner_parity(oracle_mod, log2p):
    _case(f"curved setup_preconditioner log2p={log2p}")
    from hpmp import workloads as W
    m, crds, G, O = _ring(oracle_mod, log2p, seed=4)
    st = W.noisy_freestream(m, 1 << log2p, seed=5, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    assert G.setup_preconditioner() == 0 and O.setup_preconditioner() == 0
    gv, gs, gt = G.get_preconditioner()
    ov, os_, ot = O.get_prcn()
    assert rel_err([gt], [ot]) < 1e-13
    assert rel_err([gv], [ov]) < 1e-13
    assert rel_err([gs], [os_]) < 1e-13


def test_curved_divergence_free_field_continuity_vanishes():
    _case("curved p=2 divergence-free linear field continuity residual")
    from hpmp import curved as CV
    from hpmp.tri_hp import TriHpIns
    m = CV.annulus(4, 12, jitter=0.3, seed=5)
    crds = CV.side_modes(m, m.ebdry_seg, _onb(), 2)
    st = linear_field_state(m, crds, np.array([[0.3, 1.0], [-0.7, -0.3]]), 2, m.ebdry_seg)
    G = TriHpIns(m, log2p=1, mu=1e-2)
    G.set_curved(m.ebdry_seg, crds)
    G.set_state(*st)
    G.rsdl(1)
    v, s, i = G.get_residual()
    r = max(np.abs(v[:, 2]).max(), np.abs(s[:, :, 2]).max())
    print("continuity residual", r, flush=True)
    assert r < 1e-14


def test_curved_pmg_cycle_parity(oracle_mod):
    _case("curved p-multigrid cycle x3 log2p=2 mglvls=3")
    from hpmp import workloads as W
    m, crds, G, O = _ring(oracle_mod, 2, seed=6)
    st = W.noisy_freestream(m, 4, seed=12, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    for _ in range(3):
        eg = G.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
        eo = O.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
        assert abs(eg - eo) <= 1e-10 * abs(eo)
    assert rel_err(G.get_state(), O.get_state()) < 1e-11


@pytest.mark.parametrize("log2p", [1, 2])
def test_curved_fp32_residual_matches_oracle(oracle_mod, log2p):
    _case(f"curved FP32 residual vs oracle log2p={log2p}")
    from hpmp import workloads as W
    m, crds, G, O = _ring(oracle_mod, log2p, seed=7)
    st = W.noisy_freestream(m, 1 << log2p, seed=8, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    r32 = G.residual_fp32()
    O.rsdl()
    err = rel_err(r32, O.get_res())
    print("rel err", err, flush=True)
    assert err < 5e-5


def test_set_curved_rejects_bad_input():
    _case("hpg_set_curved argument checks")
    from hpmp import curved as CV
    from hpmp.tri_hp import TriHpIns
    m = CV.annulus(3, 12)
    G = TriHpIns(m, log2p=1, mu=1e-2)
    interior = int(np.flatnonzero(~m.boundary_segment_mask())[0])
    with pytest.raises(RuntimeError):
        G.set_curved([interior], np.zeros((1, 1, 2)))
    b = int(m.ebdry_seg[0])
    with pytest.raises(RuntimeError):
        G.set_curved([b, b], np.zeros((2, 1, 2)))
    with pytest.raises(RuntimeError):
        G.set_curved([b], np.full((1, 1, 2), np.nan))
    G.set_curved(m.ebdry_seg, CV.side_modes(m, m.ebdry_seg, _onb(), 2))
    with pytest.raises(RuntimeError):
        G.element_jacobian(np.arange(4))
    G.set_curved([], np.zeros((0, 1, 2)))
    assert G.element_jacobian(np.arange(4)).shape[0] == 4


def test_taylor_couette_converges_with_curved_sides(oracle_mod):
    """p = 4 on a 4 x 24 ring: the converged vertex velocities are within 2e-6 of the
    exact Couette flow with curved sides, more than 1000 times closer than with straight
    sides (the oracle gives 3.8e-7 and 5.9e-3); the oracle's FP64 residual vanishes (1e-10 of the initial one) at the GPU's
    converged state; the FP32 defect correction reaches the same solution."""
    _case("Taylor-Couette 4x24 p=4 mu=0.1 curved vs straight, FP64 and mixed")
    from hpmp import _lib, curved as CV
    from hpmp.tri_hp import TriHpIns, rel_diff_mod_pressure
    m = CV.annulus(4, 24)
    p = 4
    crds = CV.side_modes(m, m.ebdry_seg, _onb(), p)
    bv = m.boundary_vertex_mask()
    errs, states = {}, {}
    for tag, curved, prec in (("curved", True, None), ("straight", False, None), ("curved-mixed", True, "mixed")):
        exact, st = couette_problem(m, p, crds, curved)
        G = TriHpIns(m, log2p=2, mu=0.1)
        if curved:
            G.set_curved(m.ebdry_seg, crds)
        if prec:
            G.set_precision(_lib.HPG_FP32_DEFECT, ninner=1)
        G.set_state(*st)
        G.rsdl()
        r0 = G.resnorm()
        for it in range(3000):
            G.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
            if it % 20 == 19:
                G.rsdl()
                if G.resnorm() < 1e-11 * r0:
                    break
        G.rsdl()
        rg = G.resnorm()
        sol = G.get_state()
        ex = exact(m.pnts)
        errs[tag] = np.abs(sol[0][~bv, :2] - ex[~bv, :2]).max()
        states[tag] = sol
        print(tag, "cycles", it + 1, "judge", rg, "from", r0, "vertex velocity error", errs[tag], flush=True)
        assert rg < 1e-10 * r0
        if tag == "curved":
            O = oracle_mod.TriHpIns(m, log2p=2, mu=0.1)
            O.set_curved(m.ebdry_seg, crds)
            O.set_state(*sol)
            O.rsdl()
            rv, rs, ri = O.get_res()
            bs = m.boundary_segment_mask()
            rv[bv, :2] = 0.0
            rs[bs, :, :2] = 0.0
            ro = max(np.abs(rv).max(), np.abs(rs).max(), np.abs(ri).max())
            print("oracle FP64 residual at the GPU state", ro, flush=True)
            assert ro < 1e-10 * r0
        G.close()
    assert errs["curved"] < 2e-6 and errs["curved"] * 1000 < errs["straight"]
    assert rel_diff_mod_pressure(states["curved-mixed"], states["curved"]) < 1e-9
