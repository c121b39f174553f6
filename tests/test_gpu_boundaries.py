"""Boundary types of tri_hp_ins on the GPU (hpg_set_edge_flux and the essential masks)
against the oracle: inflow (u, v essential), outflow (nothing essential), symmetry (v
essential) with the generic flux, and the euler and characteristic fluxes
(tri_hp/ins/bdry_ins.h:34-113, tri_hp/ins/bdry.cpp:245-320)."""
import numpy as np
import pytest

from conftest import rel_err
from test_oracle_boundaries import typed_square

pytestmark = pytest.mark.gpu
RSDL_TOL = 1e-12
IBC = (0.8, 0.3, 0.25)


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _pair(oracle_mod, n, log2p, types, seed=1, mu=1e-2):
    from hpmp.tri_hp import TriHpIns
    m, _ = typed_square(n, seed=seed)
    G = TriHpIns(m, log2p=log2p, mu=mu)
    O = oracle_mod.TriHpIns(m, log2p=log2p, mu=mu)
    G.set_edge_flux(types, IBC)
    O.set_edge_flux(types, IBC)
    return m, G, O


@pytest.mark.parametrize("log2p", [0, 1, 2])
@pytest.mark.parametrize("types", [[0, 0, 2, 0], [0, 1, 2, 1]])
def test_boundary_types_rsdl_parity(oracle_mod, log2p, types):
    _case(f"boundary types {types} rsdl log2p={log2p}")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 8, log2p, types, seed=3)
    st = W.noisy_freestream(m, 1 << log2p, seed=11, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    for s in (0, 1, 2):
        G.rsdl(s); O.rsdl(s)
        assert rel_err(G.get_residual(), O.get_res()) < RSDL_TOL
    G.rsdl(); O.rsdl()
    assert rel_err(G.get_residual(), O.get_res()) < RSDL_TOL


@pytest.mark.parametrize("log2p", [1, 2])
def test_boundary_types_fp32_residual(oracle_mod, log2p):
    _case(f"boundary types FP32 residual log2p={log2p}")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 8, log2p, [0, 1, 2, 0], seed=4)
    st = W.noisy_freestream(m, 1 << log2p, seed=12, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    r32 = G.residual_fp32()
    O.rsdl()
    err = rel_err(r32, O.get_res())
    print("rel err", err, flush=True)
    assert err < 5e-5


def test_boundary_types_converged_solution_matches(oracle_mod):
    """Inflow left, outflow right, symmetry below, characteristic far field above: from a
    perturbed freestream both relaxations reach the same steady state (p = 2, 6 x 6)."""
    _case("boundary types converged steady solution log2p=1 n=6")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 6, 1, [0, 0, 2, 0], seed=15, mu=5e-2)
    st = W.noisy_freestream(m, 2, seed=16, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    G.rsdl()
    r0 = G.resnorm()
    for it in range(6000):
        G.setup_preconditioner(); G.update()
        if it % 50 == 49:
            G.rsdl()
            if G.resnorm() < 1e-11 * r0:
                break
    for _ in range(it + 1):
        O.setup_preconditioner(); O.update()
    G.rsdl()
    print("iterations", it + 1, "judge", G.resnorm(), "from", r0, flush=True)
    assert G.resnorm() < 1e-10 * r0
    assert rel_err(G.get_state(), O.get_state()) < 1e-10


def test_boundary_types_pmg_cycle_parity(oracle_mod):
    _case("boundary types p-multigrid cycle x3 log2p=2")
    from hpmp import workloads as W
    m, G, O = _pair(oracle_mod, 8, 2, [0, 1, 2, 0], seed=10)
    st = W.noisy_freestream(m, 4, seed=12, scale=0.3)
    G.set_state(*st); O.set_state(*st)
    for _ in range(3):
        eg = G.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
        eo = O.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
        assert abs(eg - eo) <= 1e-10 * abs(eo)
    assert rel_err(G.get_state(), O.get_state()) < 1e-11


def test_set_edge_flux_rejects_bad_input():
    _case("hpg_set_edge_flux argument checks")
    from hpmp.tri_hp import TriHpIns
    m, _ = typed_square(4)
    G = TriHpIns(m, log2p=1, mu=1e-2)
    with pytest.raises(RuntimeError):
        G.set_edge_flux([0, 0, 3, 0], IBC)
    with pytest.raises(RuntimeError):
        G.set_edge_flux([0, 0, 2], IBC)
    with pytest.raises(RuntimeError):
        G.set_edge_flux([0, 0, 2, 0], (0.0, np.nan, 0.0))
    G.set_edge_flux([0, 0, 2, 0], IBC)
