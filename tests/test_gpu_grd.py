"""GPU parity on the unstructured meshes of the reference's own test decks, read from
their .grd files (tests/golden/grd): setup_preconditioner + update + rsdl against the
oracle to 1e-12 relative (a NaN or infinity anywhere fails), ins at p = 4 and cns
at p = 2."""
import os

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(__file__), "golden", "grd")
INFLOW = 0b1110


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


@pytest.mark.parametrize("name", ["tim.1", "channel", "square4"])
def test_ins_on_reference_mesh(oracle_mod, name):
    _case(f"tri ins p=4 on {name}.grd")
    from hpmp import grd, workloads as W
    from hpmp.tri_hp import TriHpIns
    m = grd.read_grd(os.path.join(HERE, name + ".grd"))
    st = W.noisy_freestream(m, 4, seed=7, scale=0.3)
    G = TriHpIns(m, log2p=2, mu=1e-2)
    O = oracle_mod.TriHpIns(m, log2p=2, mu=1e-2)
    for X in (G, O):
        X.set_state(*st)
        X.setup_preconditioner()
        X.update()
        X.rsdl()
    err = rel_err(G.get_residual(), O.get_res())
    print("rel err", err, flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("name", ["tim.1", "channel"])
def test_cns_on_reference_mesh(oracle_mod, name):
    _case(f"tri cns p=2 on {name}.grd")
    from hpmp import grd, workloads as W
    from hpmp.tri_hp import TriHpCns
    m = grd.read_grd(os.path.join(HERE, name + ".grd"), essential=INFLOW)
    st, bc = W.cns_gaussian_state(m, 2, seed=5, amp=0.05, width=0.2, noise=2e-2)
    # the reference's explicit cns relaxation stays admissible from this state only
    # at a small CFL (its cns decks drive PETSc instead)
    cfl = (0.01, 0.01, 0.01)
    G = TriHpCns(m, log2p=1, mu=1e-2, bc=bc, cfl=cfl)
    O = oracle_mod.TriHpCns(m, log2p=1, mu=1e-2, bc=bc, cfl=cfl)
    for X in (G, O):
        X.set_state(*st)
        assert X.setup_preconditioner() == 0
        X.update()
        X.rsdl()
    err = rel_err(G.get_residual(), O.get_res())
    print("rel err", err, flush=True)
    assert err < 1e-12


def test_h_cycle_on_reference_mesh_with_coarse_meshes(oracle_mod):
    """h+p multigrid on tim.1.grd with the coarse meshes and connection tables of
    hpmp.coarsen (unstructured coarsening, tri_mesh/coarsen.cpp, tri_mesh::connect):
    the fused GPU cycle (captured as a graph from its second call) against the oracle's
    cycle on the same hierarchy, ins p = 4, three W cycles, to 1e-11."""
    _case("h+p W cycle on tim.1.grd with unstructured coarse meshes, ins p=4")
    from hpmp import grd, multigrid as MG, workloads as W
    from hpmp.coarsen import hierarchy
    from hpmp.tri_hp import TriHpIns
    m = grd.read_grd(os.path.join(HERE, "tim.1.grd"))
    st = W.noisy_freestream(m, 4, seed=7, scale=0.3)
    meshes, links = hierarchy(m, 3)
    assert len(meshes) == 3
    G = TriHpIns(m, log2p=2, mu=1e-2)
    O = oracle_mod.TriHpIns(m, log2p=2, mu=1e-2)
    gg = MG.build_stack(TriHpIns, G, meshes, links, mu=1e-2)
    og = MG.build_stack(oracle_mod.TriHpIns, O, meshes, links, mu=1e-2)
    G.set_state(*st)
    O.set_state(*st)
    for k in range(3):
        eg = MG.cycle(gg, 2, 2, 1, 1)
        eo = MG.cycle(og, 2, 2, 1, 1)
        print("cycle", k, "maxres gpu", eg, "oracle", eo, flush=True)
        assert abs(eg - eo) <= 1e-10 * abs(eo)
    err = rel_err(G.get_state(), O.get_state())
    print("state rel err", err, flush=True)
    assert err < 1e-11
