"""GPU parity: libhpgpu.so (through the C-ABI) against the CPU oracle.

Bar: residuals to 1e-12 relative in max norm, converged solutions to
1e-10.  Every test prints the case it is about to run before touching the GPU.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
RSDL_TOL = 1e-12


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _pair(oracle_mod, n, log2p, jitter=0.2, seed=1, mu=1e-2, essential=0b011, **kw):
    from hpmp import mesh as M
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(n, jitter, seed, essential_mask=essential)
    return m, TriHpIns(m, log2p=log2p, mu=mu, **kw), oracle_mod.TriHpIns(m, log2p=log2p, mu=mu, **kw)


def test_mfma_f64_lane_layout():
    _case("mfma f64 16x16x4 layout")
    from hpmp import _lib
    L = _lib.load()
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 4)); B = rng.standard_normal((4, 16))
    out = np.zeros(256)
    assert L.hpg_selftest_mfma(_lib.ptr(np.ascontiguousarray(A)), _lib.ptr(np.ascontiguousarray(B)), _lib.ptr(out)) == 0
    D = A @ B
    exp = np.array([[D[4 * i + l // 16, l % 16] for i in range(4)] for l in range(64)]).ravel()
    assert np.abs(out - exp).max() < 1e-13


@pytest.mark.parametrize("log2p", [0, 1, 2])
@pytest.mark.parametrize("stage", [5, 0, 1, 2])
def test_rsdl_parity(oracle_mod, log2p, stage):
    _case(f"rsdl log2p={log2p} stage={stage}")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 10, log2p, seed=3)
    st = W.noisy_freestream(m, 1 << log2p, seed=11)
    G.set_state(*st); O.set_state(*st)
    # stages > 0 read res_r left by the earlier stages: run the same sequence on both
    for s in range(stage + 1) if stage < 5 else [5]:
        G.rsdl(s); O.rsdl(s)
    assert rel_err(G.get_residual(), O.get_res()) < RSDL_TOL


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_setup_preconditioner_parity(oracle_mod, log2p):
    _case(f"setup_preconditioner log2p={log2p}")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 9, log2p, seed=4)
    st = W.noisy_freestream(m, 1 << log2p, seed=5)
    G.set_state(*st); O.set_state(*st)
    assert G.setup_preconditioner() == 0 and O.setup_preconditioner() == 0
    gv, gs, gt = G.get_preconditioner()
    ov, os_, ot = O.get_prcn()
    assert rel_err([gt], [ot]) < 1e-13
    assert rel_err([gv], [ov]) < 1e-13
    if log2p:
        assert rel_err([gs], [os_]) < 1e-13


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_minvrt_parity(oracle_mod, log2p):
    _case(f"minvrt log2p={log2p}")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 9, log2p, seed=6)
    st = W.noisy_freestream(m, 1 << log2p, seed=7)
    G.set_state(*st); O.set_state(*st)
    G.setup_preconditioner(); O.setup_preconditioner()
    O.rsdl()
    r = O.get_res()
    G.set_residual(*r)
    O.minvrt(); G.minvrt()
    assert rel_err(G.get_residual(), O.get_res()) < RSDL_TOL


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_update_parity(oracle_mod, log2p):
    _case(f"update x3 log2p={log2p}")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 8, log2p, seed=8)
    st = W.noisy_freestream(m, 1 << log2p, seed=9)
    G.set_state(*st); O.set_state(*st)
    for _ in range(3):
        G.setup_preconditioner(); O.setup_preconditioner()
        G.update(); O.update()
    assert rel_err(G.get_state(), O.get_state()) < 1e-12
    assert rel_err(G.get_residual(), O.get_res()) < 1e-10


def test_pmg_cycle_parity(oracle_mod):
    _case("p-multigrid cycle x3 log2p=2 mglvls=3")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 8, 2, seed=10)
    st = W.noisy_freestream(m, 4, seed=12)
    G.set_state(*st); O.set_state(*st)
    for _ in range(3):
        eg = G.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
        eo = O.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
        assert abs(eg - eo) <= 1e-10 * abs(eo)
    assert G.log2p == 2
    assert rel_err(G.get_state(), O.get_state()) < 1e-11


def test_maxres_and_resnorm(oracle_mod):
    _case("maxres/resnorm log2p=2")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 8, 2, seed=13)
    st = W.noisy_freestream(m, 4, seed=14)
    G.set_state(*st); O.set_state(*st)
    G.rsdl(); O.rsdl()
    eg, pvg = G.maxres()
    eo, pvo = O.maxres()
    assert abs(eg - eo) <= 1e-12 * eo and np.abs(pvg - pvo).max() <= 1e-12 * eo
    rv, rs, ri = O.get_res()
    bv = m.boundary_vertex_mask(); bs = m.boundary_segment_mask()
    rv = rv.copy(); rs = rs.copy()
    rv[bv, :2] = 0.0; rs[bs, :, :2] = 0.0
    ref = max(np.abs(rv).max(), np.abs(rs).max(), np.abs(ri).max())
    assert abs(G.resnorm() - ref) <= 1e-12 * ref


def test_rsdl_is_deterministic():
    _case("rsdl run-to-run bitwise log2p=2")
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(24, 0.2, 2)
    G = TriHpIns(m, log2p=2, mu=1e-3)
    G.set_state(*W.fourier_state(m, 4, seed=1))
    G.rsdl()
    a = G.get_residual()
    for _ in range(3):
        G.rsdl()
        b = G.get_residual()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_converged_solution_matches(oracle_mod):
    _case("converged steady solution log2p=1 n=6")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 6, 1, seed=15, mu=5e-2)
    st = W.noisy_freestream(m, 2, seed=16, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    G.rsdl()
    r0 = G.resnorm()
    for it in range(4000):
        G.setup_preconditioner(); G.update()
        if it % 50 == 49:
            G.rsdl()
            if G.resnorm() < 1e-11 * r0:
                break
    for _ in range(it + 1):
        O.setup_preconditioner(); O.update()
    G.rsdl()
    assert G.resnorm() < 1e-10 * r0
    assert rel_err(G.get_state(), O.get_state()) < 1e-10


def test_full_size_free_stream_and_parity(oracle_mod):
    """512 x 512 squares, p = 4 (524 288 elements): free stream gives a zero residual
    and one full residual of the jittered Fourier state matches the oracle."""
    _case("full size 512x512 p=4 rsdl")
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(512, 0.2, 1)
    G = TriHpIns(m, log2p=2, mu=1e-3)
    G.set_state(*W.freestream(m, 4))
    G.rsdl()
    for a in G.get_residual():
        assert np.abs(a).max() < 1e-13
    st = W.fourier_state(m, 4, seed=2)
    G.set_state(*st)
    G.rsdl()
    O = oracle_mod.TriHpIns(m, log2p=2, mu=1e-3)
    O.set_state(*st)
    O.rsdl()
    assert rel_err(G.get_residual(), O.get_res()) < RSDL_TOL
