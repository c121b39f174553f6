"""input_map: the reference's key: value decks with muParser formula values
(input_map/input_map.cpp:80-227) and the ins edge-boundary types that map to essential
masks (tri_hp/allocate.cpp:182-186, tri_hp/ins/getnewbdry.cpp:95-113).

Pins: the reference's own convergence deck (tests/golden/decks, copied unchanged) reads
back its cycle controls and its formula list ``cfls: C*2.5 C*1.5 C*1.0``; formulas
against Python's own arithmetic on the same numbers; chained variables; errors for
self-reference, undefined names and anything outside the formula grammar.
"""
import math
import os

import pytest

from hpmp.input_map import FormulaError, InputMap, evaluate, ins_boundary_essential

DECK = os.path.join(os.path.dirname(__file__), "golden", "decks", "ins_convergence_run_master.inpt")


def test_reference_deck_values():
    m = InputMap.from_file(DECK)
    assert m.getlist("cfls", (0.0,), 3) == [2.5, 1.5, 1.0]
    assert (m.geti("vwcycle", 1), m.geti("itercrsn", 1), m.geti("iterrfne", 0)) == (2, 2, 2)
    assert m.geti("ngrid", 1) == 2 and m.geti("ncycle", 0) == 85
    assert m.getf("rho", 0.0) == 1.0 and m.getf("mu", 1.0) == 0.0
    assert m.getf("relative_tolerance", 0.0) == 1.0e-8
    # absent keys take the default
    assert m.getf("not_there", 3.5) == 3.5 and m.getlist("alpha", (0.25, 1.0), 2) == [0.25, 1.0]


@pytest.mark.parametrize("expr,value", [
    ("2^3", 8.0), ("2^3^2", 2.0 ** 9), ("-3+4*2", 5.0), ("(1+2)*(3-5)/4", -1.5), ("1.5e-3*2", 3.0e-3),
    ("sqrt(2)*sin(_pi/4)", math.sqrt(2) * math.sin(math.pi / 4)), ("exp(1)-_e", 0.0), ("ln(_e^2)", 2.0),
    ("log10(1000)", 3.0), ("erf(0.5)+erfc(0.5)", 1.0), ("abs(-2)+sign(-3)+rint(2.6)", 4.0),
    ("min(3,1,2)+max(3,1,2)+sum(1,2,3)+avg(2,4)", 13.0), ("2<3", 1.0), ("2>=3", 0.0),
    ("1<2 && 2<1", 0.0), ("1<2 || 2<1", 1.0), ("3>2 ? 10 : 20", 10.0), ("1>2 ? 10 : 2<3 ? 30 : 40", 30.0),
])
def test_formula_values(expr, value):
    assert evaluate(expr) == pytest.approx(value, rel=1e-15, abs=1e-15)


def test_chained_variables_and_block_prefix():
    m = InputMap.from_text("a: 2\nb: a*3\nc: b^2 - a\nb0_mu: c/1000\nmu: 7\nlst: a b c a*b\n")
    assert m.value("c") == 34.0
    assert m.getf("mu", 0.0) == pytest.approx(0.034, rel=1e-15)  # b0_mu before mu
    assert m.getf("mu", 0.0, prefix="b1") == 7.0
    assert m.getlist("lst", (), 4) == [2.0, 6.0, 34.0, 12.0]
    assert m.geti("c", 0) == 34


@pytest.mark.parametrize("text,key", [
    ("x: x+1", "x"), ("x: y*2\ny: x", "x"), ("x: z+1", "x"), ("x: __import__('os')", "x"),
    ("x: (1).real", "x"), ("x: [1][0]", "x"), ("x: 'a'", "x"), ("x: lambda: 1", "x"), ("x: 1 ? 2", "x"),
])
def test_formula_errors(text, key):
    with pytest.raises(FormulaError):
        InputMap.from_text(text).value(key)


def test_boundary_types_to_essential_masks():
    deck = InputMap.from_text("b0_s1_hp_type: inflow\nb0_s2_hp_type: outflow\nb0_s3_hp_type: symmetry\n"
                              "b0_s3_dir: 1\nb0_s4_hp_type: symmetry\n")
    assert ins_boundary_essential(deck, [1, 2, 3, 4]) == {1: 0b011, 2: 0, 3: 0b010, 4: 0b001}
    ref = InputMap.from_file(DECK)
    assert ins_boundary_essential(ref, [2, 4]) == {2: 0, 4: 0b011}
    with pytest.raises(ValueError, match="periodic"):
        ins_boundary_essential(ref, [1])   # b0_s1_type: prdc
    with pytest.raises(ValueError, match="plain"):
        ins_boundary_essential(ref, [3])   # no hp_type: the reference's default "plain"
    with pytest.raises(ValueError, match="characteristic"):
        ins_boundary_essential(InputMap.from_text("b0_s1_hp_type: characteristic"), [1])
