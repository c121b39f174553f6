"""GPU parity of the tet_hp_ins path (hpg_tet_* through the C-ABI) against the CPU
oracle (oracle/otet_hp.cpp).  Residuals to 1e-12 relative in max norm; the
converged solution to 1e-10 (modulo the constant-pressure null mode)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
RSDL_TOL = 1e-12


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _state(m, em0, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    v = np.zeros((m.npnt, 4))
    v[:, 0], v[:, 1], v[:, 2], v[:, 3] = 1.0, 0.3, -0.2, 0.5
    v2 = v + amp * rng.standard_normal(v.shape)
    e2 = amp * rng.standard_normal((m.nseg, em0, 4))
    v2[m.boundary_vertex_mask(), :3] = v[m.boundary_vertex_mask(), :3]
    e2[m.boundary_edge_mask(), :, :3] = 0.0
    return v2, e2


def _pair(n, log2p, mu=1e-2, seed=3, jitter=0.2):
    import oracle_tet as OT
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpIns
    m = TM.unit_cube(n, jitter, seed)
    st = _state(m, (1 << log2p) - 1, seed + 10)
    G = TetHpIns(m, log2p=log2p, mu=mu)
    O = OT.TetHpIns(m, log2p=log2p, mu=mu)
    G.set_state(*st)
    O.set_state(*st)
    return m, st, G, O


@pytest.mark.parametrize("log2p", [0, 1])
def test_tet_setup_preconditioner_parity(oracle_mod, log2p):
    _case(f"tet setup_preconditioner log2p={log2p}")
    m, st, G, O = _pair(4, log2p)
    assert G.setup_preconditioner() == 0 and O.setup_preconditioner() == 0
    g, o = G.get_preconditioner(), O.get_prcn()
    names = ("vprcn", "eprcn", "iprcn", "tau")
    for k in range(4):
        if k == 1 and log2p == 0:
            continue
        a, b = g[k], o[k]
        if k == 3:
            a, b = a[:, [0, 3]], b[:, [0, 3]]
        err = rel_err([a], [b])
        print(names[k], "rel err", err, flush=True)
        assert err < 1e-13


@pytest.mark.parametrize("log2p", [0, 1])
@pytest.mark.parametrize("stage", [5, 0, 1, 2])
def test_tet_rsdl_parity(oracle_mod, log2p, stage):
    _case(f"tet rsdl log2p={log2p} stage={stage}")
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
def test_tet_minvrt_parity(oracle_mod, log2p):
    _case(f"tet minvrt log2p={log2p}")
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
def test_tet_update_parity(oracle_mod, log2p):
    _case(f"tet update x3 log2p={log2p}")
    m, st, G, O = _pair(4, log2p, seed=9)
    for _ in range(3):
        G.setup_preconditioner()
        O.setup_preconditioner()
        G.update()
        O.update()
    err = rel_err(G.get_state(), O.get_state())
    print("state rel err", err, flush=True)
    assert err < 1e-12


def test_tet_pmg_cycle_parity(oracle_mod):
    _case("tet p-multigrid cycles 2 -> 1")
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


def test_tet_rsdl_is_deterministic():
    _case("tet rsdl determinism 6^3 p=2")
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpIns
    m = TM.unit_cube(6, 0.2, 4)
    G = TetHpIns(m, log2p=1, mu=1e-2)
    G.set_state(*_state(m, 1, 5))
    G.setup_preconditioner()
    G.rsdl()
    a = G.get_residual()
    G.rsdl()
    b = G.get_residual()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_tet_uniform_stream_full_size():
    _case("tet uniform stream 64^3 p=2 (1 572 864 tets)")
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpIns
    m = TM.unit_cube(64)
    G = TetHpIns(m, log2p=1, mu=1e-2)
    v = np.zeros((m.npnt, 4))
    v[:, 0], v[:, 1], v[:, 2], v[:, 3] = 1.0, 0.3, -0.2, 0.5
    G.set_state(v, np.zeros((m.nseg, 1, 4)))
    assert G.setup_preconditioner() == 0
    G.rsdl()
    rv, re = G.get_residual()
    bv, be = m.boundary_vertex_mask(), m.boundary_edge_mask()
    mx = max(np.abs(rv[~bv]).max(), np.abs(re[~be]).max(), np.abs(rv[:, 3]).max(), np.abs(re[:, :, 3]).max())
    print("max |res| (interior, continuity)", mx, flush=True)
    assert mx < 1e-12


def test_tet_converged_solution_matches(oracle_mod):
    _case("tet converged solution 3^3 p=2")
    m, st, G, O = _pair(3, 1, mu=5e-2, seed=2)
    bv, be = m.boundary_vertex_mask(), m.boundary_edge_mask()

    def run(X, get_res):
        for it in range(3000):
            X.setup_preconditioner()
            X.update()
            if it % 100 == 99:
                X.rsdl()
                rv, re = get_res()
                rv[bv, :3] = 0
                re[be, :, :3] = 0
                if max(np.abs(rv).max(), np.abs(re).max()) < 1e-13:
                    return it + 1
        return it + 1

    ng = run(G, G.get_residual)
    no = run(O, O.get_res)
    a, b = G.get_state(), O.get_state()
    dv = a[0] - b[0]
    dv[:, 3] -= dv[:, 3].mean()
    err = max(np.abs(dv).max(), np.abs(a[1] - b[1]).max()) / np.abs(b[0]).max()
    print("iterations gpu", ng, "oracle", no, "solution rel diff (mod pressure)", err, flush=True)
    assert err < 1e-10
