"""GPU parity of the tet_hp_cns path (hpg_tet_* with HPG_PHYSICS_CNS through the C-ABI)
against the CPU oracle (oracle/otet_cns.cpp).  Residuals to 1e-12 relative in max norm;
the converged solution to 1e-10."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
RSDL_TOL = 1e-12
STATE = np.array([1.0, 0.3, 0.1, -0.05, 1.0])   # p, u, v, w, RT: Mach ~0.27 at gamma 1.4, R 1
KW = dict(gamma=1.4, prandtl=0.72, R=1.0)


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _state(m, em0, seed, amp=0.02):
    rng = np.random.default_rng(seed)
    v = np.tile(STATE, (m.npnt, 1)) + amp * rng.standard_normal((m.npnt, 5))
    e = amp * rng.standard_normal((m.nseg, em0, 5))
    v[m.boundary_vertex_mask(), 1:] = STATE[1:]
    e[m.boundary_edge_mask(), :, 1:] = 0.0
    return v, e


def _pair(n, log2p, mu=1e-2, seed=3, jitter=0.2):
    import oracle_tet as OT
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpCns
    m = TM.unit_cube(n, jitter, seed)
    st = _state(m, (1 << log2p) - 1, seed + 10)
    G = TetHpCns(m, log2p=log2p, mu=mu, bc=tuple(STATE[1:]), **KW)
    O = OT.TetHpCns(m, log2p=log2p, mu=mu, bc=(0.0, *STATE[1:]), **KW)
    G.set_state(*st)
    O.set_state(*st)
    return m, st, G, O


@pytest.mark.parametrize("log2p", [0, 1])
def test_tetc_setup_preconditioner_parity(oracle_mod, log2p):
    _case(f"tet cns setup_preconditioner log2p={log2p}")
    m, st, G, O = _pair(4, log2p)
    assert G.setup_preconditioner() == 0 and O.setup_preconditioner() == 0
    gv, ge, gi, gt = G.get_preconditioner()
    gvpc, gepc = G.get_cns_blocks()
    ov, oe, ot, ovpc, oepc, _ = O.get_prcn()
    checks = [("tau", gt, ot), ("vprcn", gv, ov), ("vpreconditioner", gvpc, ovpc)]
    if log2p:
        checks += [("eprcn", ge, oe), ("epreconditioner", gepc, oepc)]
    for name, a, b in checks:
        err = rel_err([a], [b])
        print(name, "rel err", err, flush=True)
        assert err < 1e-13


@pytest.mark.parametrize("log2p", [0, 1])
@pytest.mark.parametrize("stage", [5, 0, 1, 2])
def test_tetc_rsdl_parity(oracle_mod, log2p, stage):
    _case(f"tet cns rsdl log2p={log2p} stage={stage}")
    m, st, G, O = _pair(4, log2p, seed=stage + 1)
    G.setup_preconditioner()
    O.setup_preconditioner()
    for s in range(stage + 1) if stage < 5 else [5]:
        G.rsdl(s)
        O.rsdl(s)
    err = rel_err(G.get_residual(), O.get_res())
    print("rel err", err, flush=True)
    assert err < RSDL_TOL


@pytest.mark.parametrize("log2p", [0, 1])
def test_tetc_minvrt_parity(oracle_mod, log2p):
    _case(f"tet cns minvrt log2p={log2p}")
    m, st, G, O = _pair(4, log2p, seed=7)
    G.setup_preconditioner()
    O.setup_preconditioner()
    rng = np.random.default_rng(8)
    res = [rng.standard_normal(a.shape) for a in G.zeros()]
    G.set_residual(*res)
    O.set_res(*res)
    G.minvrt()
    O.minvrt()
    err = rel_err(G.get_residual(), O.get_res())
    print("rel err", err, flush=True)
    assert err < 1e-13


@pytest.mark.parametrize("log2p", [0, 1])
def test_tetc_update_parity(oracle_mod, log2p):
    _case(f"tet cns update x3 log2p={log2p}")
    m, st, G, O = _pair(4, log2p, seed=9)
    for _ in range(3):
        assert G.setup_preconditioner() == 0
        O.setup_preconditioner()
        G.update()
        O.update()
    err = rel_err(G.get_state(), O.get_state())
    print("state rel err", err, flush=True)
    assert err < 1e-12


def test_tetc_pmg_cycle_parity(oracle_mod):
    _case("tet cns p-multigrid cycles 2 -> 1")
    m, st, G, O = _pair(4, 1, seed=11)
    for _ in range(3):
        for X in (G, O):
            X.setup_preconditioner()
            X.update()
            X.rsdl()
            X.mg_restrict()
            X.setup_preconditioner()
            X.update()
            X.mg_prolongate()
    err = rel_err(G.get_state(), O.get_state())
    print("state rel err", err, flush=True)
    assert err < 1e-11


def test_tetc_rsdl_is_deterministic():
    _case("tet cns rsdl determinism 6^3 p=2")
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpCns
    m = TM.unit_cube(6, 0.2, 4)
    G = TetHpCns(m, log2p=1, mu=1e-2, bc=tuple(STATE[1:]), **KW)
    G.set_state(*_state(m, 1, 5))
    G.setup_preconditioner()
    G.rsdl()
    a = G.get_residual()
    G.rsdl()
    b = G.get_residual()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_tetc_uniform_stream_full_size():
    _case("tet cns uniform stream 64^3 p=2 (1 572 864 tets)")
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpCns
    m = TM.unit_cube(64)
    G = TetHpCns(m, log2p=1, mu=1e-2, bc=tuple(STATE[1:]), **KW)
    G.set_state(np.tile(STATE, (m.npnt, 1)), np.zeros((m.nseg, 1, 5)))
    assert G.setup_preconditioner() == 0
    G.rsdl()
    rv, re = G.get_residual()
    mx = max(np.abs(rv).max(), np.abs(re).max())
    print("max |res| (every equation, every mode)", mx, flush=True)
    assert mx < 1e-13


def test_tetc_nan_state_reports_setup_error():
    _case("tet cns NaN state -> setup error count")
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpCns
    m = TM.unit_cube(2)
    G = TetHpCns(m, log2p=1, mu=1e-2, bc=tuple(STATE[1:]), **KW)
    v, e = np.tile(STATE, (m.npnt, 1)), np.zeros((m.nseg, 1, 5))
    G.set_state(v, e)
    assert G.setup_preconditioner() == 0
    v[4, 2] = np.nan
    G.set_state(v, e)
    assert G.setup_preconditioner() >= 1


def test_tetc_converged_solution_matches(oracle_mod):
    _case("tet cns converged solution 3^3 p=2")
    m, st, G, O = _pair(3, 1, mu=5e-2, seed=2)
    bv, be = m.boundary_vertex_mask(), m.boundary_edge_mask()

    def run(X, get_res):
        for it in range(2000):
            assert X.setup_preconditioner() == 0
            X.update()
            if it % 100 == 99:
                X.rsdl()
                rv, re = get_res()
                rv[bv, 1:] = 0
                re[be, :, 1:] = 0
                if max(np.abs(rv).max(), np.abs(re).max()) < 1e-13:
                    return it + 1
        return it + 1

    ng = run(G, G.get_residual)
    no = run(O, O.get_res)
    a, b = G.get_state(), O.get_state()
    err = max(np.abs(a[0] - b[0]).max(), np.abs(a[1] - b[1]).max()) / np.abs(b[0]).max()
    print("iterations gpu", ng, "oracle", no, "solution rel diff", err, flush=True)
    assert err < 1e-10
