"""GPU tet element Jacobian (hpg_tet_element_jacobian, tet_hp/update.cpp:134-178)
against the oracle's, ins and cns, p = 1 and 2.  With the reference's eps (dw about
1e-5 here) the residual rounding differences are amplified to about 1e-9 of max |K|;
the tolerance is 1e-7."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _pair(cns, log2p):
    import oracle_tet as OT
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpCns, TetHpIns
    m = TM.unit_cube(3, 0.15, 5)
    rng = np.random.default_rng(21)
    if cns:
        st = np.array([1.0, 0.3, 0.1, -0.05, 1.0])
        kw = dict(gamma=1.4, prandtl=0.72, R=1.0)
        G = TetHpCns(m, log2p=log2p, mu=1e-2, bc=tuple(st[1:]), **kw)
        O = OT.TetHpCns(m, log2p=log2p, mu=1e-2, bc=(0.0, *st[1:]), **kw)
    else:
        st = np.array([1.0, 0.3, -0.2, 0.5])
        G = TetHpIns(m, log2p=log2p, mu=1e-2)
        O = OT.TetHpIns(m, log2p=log2p, mu=1e-2)
    v = np.tile(st, (m.npnt, 1)) + 0.02 * rng.standard_normal((m.npnt, st.size))
    e = 0.02 * rng.standard_normal((m.nseg, (1 << log2p) - 1, st.size))
    for X in (G, O):
        X.set_state(v, e)
        assert X.setup_preconditioner() == 0
    return m, G, O


@pytest.mark.parametrize("log2p", [0, 1])
@pytest.mark.parametrize("cns", [False, True])
def test_tet_element_jacobian_parity(oracle_mod, cns, log2p):
    _case(f"tet element_jacobian cns={cns} log2p={log2p}")
    m, G, O = _pair(cns, log2p)
    batch = np.random.default_rng(log2p).choice(m.ntet, size=12, replace=False)
    s0 = [a.copy() for a in G.get_state()]
    K = G.element_jacobian(batch)
    worst = 0.0
    for j, t in enumerate(batch):
        Ko = O.element_jacobian(int(t))
        worst = max(worst, np.abs(K[j] - Ko).max() / np.abs(Ko).max())
    print("max rel err", worst, flush=True)
    assert worst < 1e-7
    for a, b in zip(G.get_state(), s0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("cns", [False, True])
def test_tet_element_jacobian_on_coarse_p_level(oracle_mod, cns):
    """The p = 1 level of a p = 2 context (tm = 4 with em0 = 1 edge slots per edge)
    after mg_restrict, against the oracle at the same level."""
    _case(f"tet element_jacobian coarse level cns={cns}")
    m, G, O = _pair(cns, 1)
    for X in (G, O):
        X.rsdl()
        X.mg_restrict()
        assert X.log2p == 0
    batch = np.random.default_rng(7).choice(m.ntet, size=10, replace=False)
    s0 = [a.copy() for a in G.get_state()]
    K = G.element_jacobian(batch)
    worst = max(np.abs(K[j] - O.element_jacobian(int(t))).max() / np.abs(O.element_jacobian(int(t))).max()
                for j, t in enumerate(batch))
    print("max rel err", worst, flush=True)
    assert worst < 1e-7
    for a, b in zip(G.get_state(), s0):
        assert np.array_equal(a, b)


def test_tet_element_jacobian_whole_mesh_batch(oracle_mod):
    """Every element of the mesh in one batch (split into vertex-disjoint sets by the
    colouring; each set's columns evaluate that set's elements only)."""
    _case("tet element_jacobian whole mesh")
    m, G, O = _pair(False, 1)
    batch = np.arange(m.ntet)
    K = G.element_jacobian(batch)
    for t in (0, m.ntet // 2, m.ntet - 1):
        Ko = O.element_jacobian(int(t))
        assert np.abs(K[t] - Ko).max() < 1e-7 * np.abs(Ko).max()
