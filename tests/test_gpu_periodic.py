"""Periodic boundaries on the GPU (hpg_set_periodic) against the oracle's restatement of
the reference's prdc treatment (partial residuals per copy, averaged preconditioner sums
and solved values).  The GPU holds the assembled residual on every copy; states after
relaxation, multigrid cycles and convergence must agree with the oracle's, and the
GPU's residual must equal the oracle's partial residuals summed over the copies."""
import numpy as np
import pytest

from conftest import rel_err
from test_oracle_periodic import periodic_state, periodic_strip

pytestmark = pytest.mark.gpu


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _pair(oracle_mod, log2p, n=4, mu=2e-2, pairs=((0, 2),), masks=(0b000, 0b000, 0b000, 0b011), seed=3):
    from hpmp.tri_hp import TriHpIns
    m = periodic_strip(n, 1, jitter=0.3, seed=seed, masks=masks)
    G = TriHpIns(m, log2p=log2p, mu=mu)
    O = oracle_mod.TriHpIns(m, log2p=log2p, mu=mu)
    G.set_periodic(list(pairs))
    O.set_periodic(list(pairs))
    return m, G, O


def _copies(m):
    """Vertex groups and side pairs of the y-periodic strip (bottom y = 0, top y = 1)."""
    eps = 1e-9
    bot = np.flatnonzero(np.abs(m.pnts[:, 1]) < eps)
    top = np.flatnonzero(np.abs(m.pnts[:, 1] - 1.0) < eps)
    groups = [(b, top[np.argmin(np.abs(m.pnts[top, 0] - m.pnts[b, 0]))]) for b in bot]
    mid = 0.5 * (m.pnts[m.seg_pnt[:, 0]] + m.pnts[m.seg_pnt[:, 1]])
    sb = np.flatnonzero(np.abs(mid[:, 1]) < eps)
    st = np.flatnonzero(np.abs(mid[:, 1] - 1.0) < eps)
    pairs = []
    for a in sb:
        b = st[np.argmin(np.abs(mid[st, 0] - mid[a, 0]))]
        same = abs(m.pnts[m.seg_pnt[a, 0], 0] - m.pnts[m.seg_pnt[b, 0], 0]) < eps
        pairs.append((a, b, 1 if same else -1))
    return groups, pairs


def _assemble(res, groups, pairs):
    v, s, i = [np.array(x) for x in res]
    for a, b in groups:
        t = v[a] + v[b]
        v[a] = t
        v[b] = t
    for a, b, o in pairs:
        sg = np.array([1.0 if (k % 2 == 0 or o > 0) else -1.0 for k in range(s.shape[1])])[:, None]
        t = s[a] + sg * s[b]
        s[a] = t
        s[b] = sg * t
    return v, s, i


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_periodic_residual_is_the_sum_over_copies(oracle_mod, log2p):
    _case(f"periodic rsdl log2p={log2p}")
    m, G, O = _pair(oracle_mod, log2p)
    st = periodic_state(m, 1 << log2p, seed=5)
    G.set_state(*st); O.set_state(*st)
    groups, pairs = _copies(m)
    for s in (0, 1, 2):
        G.rsdl(s); O.rsdl(s)
        assert rel_err(G.get_residual(), _assemble(O.get_res(), groups, pairs)) < 1e-12
    G.rsdl(); O.rsdl()
    assert rel_err(G.get_residual(), _assemble(O.get_res(), groups, pairs)) < 1e-12


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_periodic_update_parity(oracle_mod, log2p):
    _case(f"periodic update x3 log2p={log2p}")
    m, G, O = _pair(oracle_mod, log2p, seed=4)
    st = periodic_state(m, 1 << log2p, seed=6)
    G.set_state(*st); O.set_state(*st)
    for _ in range(3):
        assert G.setup_preconditioner() == 0 and O.setup_preconditioner() == 0
        G.update(); O.update()
    assert rel_err(G.get_state(), O.get_state()) < 1e-12
    # the reference inverts the average of the copies' sums (2 / sum for a pair), the
    # GPU the assembled sum (1 / sum); the updates they give are the same
    gv, gs, gt = G.get_preconditioner()
    ov, os_, ot = [np.array(x) for x in O.get_prcn()]
    groups, pairs = _copies(m)
    for g in groups:
        ov[list(g)] /= len(g)
    for a, b, _ in pairs:
        os_[[a, b]] /= 2.0
    assert rel_err([gv], [ov]) < 1e-13
    if log2p:
        assert rel_err([gs], [os_]) < 1e-13


def test_periodic_pmg_cycle_parity(oracle_mod):
    _case("periodic p-multigrid cycle x3 log2p=2")
    m, G, O = _pair(oracle_mod, 2, seed=7)
    st = periodic_state(m, 4, seed=8)
    G.set_state(*st); O.set_state(*st)
    for _ in range(3):
        G.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
        O.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
    assert rel_err(G.get_state(), O.get_state()) < 1e-11


def test_doubly_periodic_update_parity(oracle_mod):
    """Periodic in x and y (corner vertices with four copies), pressure pinned by
    nothing: a few relaxation steps from the same state agree."""
    _case("doubly periodic update x3 log2p=1")
    m, G, O = _pair(oracle_mod, 1, pairs=((0, 2), (1, 3)), masks=(0, 0, 0, 0), seed=9)
    st = periodic_state(m, 2, seed=10)
    G.set_state(*st); O.set_state(*st)
    for _ in range(3):
        assert G.setup_preconditioner() == 0 and O.setup_preconditioner() == 0
        G.update(); O.update()
    assert rel_err(G.get_state(), O.get_state()) < 1e-12


@pytest.mark.parametrize("mixed", [False, True])
def test_periodic_converged_solution_matches(oracle_mod, mixed):
    """Inflow left, outflow right, periodic in y, p = 2: the GPU (FP64 or the FP32 defect
    correction) and the oracle reach the same steady state."""
    _case(f"periodic converged steady solution log2p=1 mixed={mixed}")
    from hpmp import _lib
    m, G, O = _pair(oracle_mod, 1, mu=5e-2, seed=11)
    st = periodic_state(m, 2, seed=12)
    G.set_state(*st); O.set_state(*st)
    if mixed:
        G.set_precision(_lib.HPG_FP32_DEFECT, ninner=1)
    G.rsdl()
    r0 = G.resnorm()
    for it in range(6000):
        G.setup_preconditioner(); G.update()
        if it % 50 == 49:
            G.rsdl()
            if G.resnorm() < 1e-12 * r0:
                break
    G.rsdl()
    rg = G.resnorm()
    # the same iterations for the FP64 form; the oracle's own convergence for the mixed one
    for _ in range(6000 if mixed else it + 1):
        O.setup_preconditioner(); O.update()
    print("iterations", it + 1, "judge", rg, "from", r0, flush=True)
    assert rg < 1e-10 * r0
    assert rel_err(G.get_state(), O.get_state()) < 1e-10


def test_set_periodic_rejects_bad_input():
    _case("hpg_set_periodic argument checks")
    from hpmp.tri_hp import TriHpIns
    m = periodic_strip(4, 1)
    G = TriHpIns(m, log2p=1, mu=1e-2)
    with pytest.raises(RuntimeError):
        G.set_periodic([(0, 1)])
    with pytest.raises(RuntimeError):
        G.set_periodic([(0, 0)])
    G.set_periodic([(0, 2)])
    with pytest.raises(RuntimeError):
        G.set_periodic([(1, 3)])
    with pytest.raises(RuntimeError):
        G.apply_jacobian(*G.zeros())
    H = TriHpIns(m, log2p=1, mu=1e-2)
    H.set_edge_flux([0, 0, 0, 0], (1.0, 0.0, 0.0))
    H.set_edge_flux([0, 2, 0, 0], (1.0, 0.0, 0.0))
    with pytest.raises(RuntimeError):
        H.set_periodic([(0, 2)])
