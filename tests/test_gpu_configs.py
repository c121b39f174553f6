"""GPU parity on the named workloads themselves, against the CPU oracle.

Workload A: the unit square split into 32 x 32 squares, each cut by the same
diagonal (2048 right triangles, no jitter), p = 2, rho = 1, mu = 1e-2, u = 1 free
stream plus N(0, 2^-k) on every coefficient of mode order k, u and v prescribed on
every side; one rsdl() and then 100 cycles of setup_preconditioner + update()
(tri_hp/nstage.cpp:12-427, tri_hp/ins/setup_preconditioner.cpp:9-310).  Residuals
to 1e-12 relative in max norm; the state after 100 cycles to 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def test_workload_a_rsdl_and_100_cycles(oracle_mod):
    _case("workload A: 32x32 p=2 mu=1e-2, rsdl + 100 cycles")
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(32, 0.0, 0)
    assert m.ntri == 2048
    st = W.noisy_freestream(m, 2, seed=2024)
    G = TriHpIns(m, log2p=1, rho=1.0, mu=1e-2)
    O = oracle_mod.TriHpIns(m, log2p=1, rho=1.0, mu=1e-2)
    G.set_state(*st); O.set_state(*st)
    G.rsdl(); O.rsdl()
    e0 = rel_err(G.get_residual(), O.get_res())
    print("rsdl rel err", e0, flush=True)
    assert e0 < 1e-12
    worst_maxres = 0.0
    for k in range(100):
        assert G.setup_preconditioner() == 0 and O.setup_preconditioner() == 0
        G.update(); O.update()
        mg, mo = G.maxres()[0], O.maxres()[0]
        worst_maxres = max(worst_maxres, abs(mg - mo) / mo)
    es = rel_err(G.get_state(), O.get_state())
    G.rsdl(); O.rsdl()
    er = rel_err(G.get_residual(), O.get_res())
    print("after 100 cycles: state rel err", es, "rsdl rel err", er, "maxres rel err", worst_maxres, flush=True)
    assert es < 1e-12 and er < 1e-12 and worst_maxres < 1e-12
