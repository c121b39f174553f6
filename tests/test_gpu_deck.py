"""The reference's Restart Test deck (Testing/tri_hp/ins/Restart Test) run through
hpmp.deck on the GPU, on an 8 x 8 periodic square written as a .grd with the deck's
boundary ids, against the oracle driven through the same setters and cycles."""
import os

import numpy as np
import pytest

from conftest import rel_err
from test_deck import RESTART
from test_oracle_periodic import periodic_strip

pytestmark = pytest.mark.gpu


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def test_restart_test_deck_runs_and_matches_the_oracle(oracle_mod, tmp_path):
    _case("Restart Test deck, 8x8 periodic square, p=1, W(2,2,2)")
    from hpmp import grd, multigrid as MG
    from hpmp.deck import Deck
    m = periodic_strip(8, 1, jitter=0.2, seed=1)
    m.ebdry_ids = [1, 2, 1, 4]
    path = str(tmp_path / "square8.grd")
    grd.write_grd(path, m)
    d = Deck(RESTART, mesh=path).build()
    hist = d.run(ncycle=10, ntstep=1)
    print("deck max residuals", ["%.3e" % e for e in hist], flush=True)
    assert all(np.isfinite(hist)) and hist[-1] < 0.5 * hist[0]
    G = d.fine
    O = oracle_mod.TriHpIns(d.mesh, log2p=0, **{k: G.params[k] for k in
                                              ("rho", "mu", "adis", "bc_pressure", "fadd", "alpha", "beta", "cfl")})
    O.set_periodic(d.pairs)
    O.set_edge_flux(d.types, d.ibc)
    O.set_state(*d.initial_state())
    for _ in range(10):
        MG.cycle([O], d.extra, d.vw, d.itercrsn, d.iterrfne, fused=False)
    assert rel_err(G.get_state(), O.get_state()) < 1e-11
    d.close()
