"""Reference input decks read by hpmp.deck (host side): the mapping of the deck's keys
onto the library's setters, and the clear errors for what has no counterpart."""
import os

import pytest

HERE = os.path.join(os.path.dirname(__file__), "golden")
RESTART = os.path.join(HERE, "decks", "ins_restart_test_run.inpt")
CONVERGENCE = os.path.join(HERE, "decks", "ins_convergence_run_master.inpt")
SQUARE1 = os.path.join(HERE, "grd", "square1.grd")


def test_restart_test_deck_mapping():
    """Testing/tri_hp/ins/Restart Test: inflow (id 4), outflow (id 2), periodic (id 1 on
    bottom and top), freestream ibc with a perturbation, W(2, 2, 2) cycles, p = 1."""
    from hpmp.deck import Deck
    d = Deck(RESTART, mesh=SQUARE1)
    m = d.mesh
    assert list(m.ebdry_ids) == [1, 2, 4, 1]
    assert d.pairs == [(0, 3)]
    # essential masks: outflow none, inflow u and v; the periodic boundaries none
    assert list(m.ebdry_essential[[1, 2]]) == [0, 0b011]
    assert d.types == [0, 0, 0, 0]
    assert d.ibc == pytest.approx((1.0, 0.0, 0.0))
    assert d.amp == pytest.approx(0.01)
    assert (d.ncycle, d.vw, d.itercrsn, d.iterrfne, d.ngrid, d.extra, d.ntstep) == (20, 2, 2, 2, 1, 0, 4)
    assert d.log2p == 0 and d.rtol == pytest.approx(1e-8)
    v, s, i = d.initial_state()
    assert v.shape == (m.npnt, 3) and abs(v[:, 0] - (1.0 + 0.01 * m.pnts[:, 0] * (1 - m.pnts[:, 0]))).max() < 1e-15


def test_convergence_deck_needs_h_levels_of_a_periodic_mesh():
    from hpmp.deck import Deck
    with pytest.raises(ValueError, match="periodic"):
        Deck(CONVERGENCE, mesh=SQUARE1)


def test_unsupported_keys_raise(tmp_path):
    from hpmp.deck import Deck
    base = open(RESTART).read()
    cases = {
        "applied_stress": base.replace("b0_s2_hp_type: outflow", "b0_s2_hp_type: applied_stress"),
        "ibc": base.replace("ibc: freestream", "ibc: symbolic"),
        "dtinv": base.replace("dtinv: 0.0", "dtinv: 10.0"),
        "only ins": base.replace("b0_type: ins", "b0_type: cns"),
    }
    for what, text in cases.items():
        p = tmp_path / f"{what.replace(' ', '_')}.inpt"
        p.write_text(text)
        with pytest.raises(ValueError, match=what):
            Deck(str(p), mesh=SQUARE1)
    with pytest.raises(FileNotFoundError):
        Deck(RESTART)   # its mesh (../../square3) is not next to the deck
