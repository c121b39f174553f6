"""The reference's text formats (hpmp.grd): .grd meshes from the reference's own test
decks (tests/golden/grd, copied unchanged by copy_grd.sh) and the text restart file."""
import os

import numpy as np
import pytest

from hpmp import grd

HERE = os.path.join(os.path.dirname(__file__), "golden", "grd")
FILES = ["square1", "square4", "channel", "tim.1"]


def _read(name, **kw):
    return grd.read_grd(os.path.join(HERE, name + ".grd"), **kw)


def test_square1_connectivity_known_answer():
    # traced by hand through tri_mesh::createsegtri (tri_mesh/setup.cpp:76-137):
    # triangle (3, 0, 1): sides (0,1) = seg 0, (1,3) = seg 4, (3,0) = seg 3, all stored
    # in the triangle's direction; triangle (1, 2, 3): (2,3) = seg 2, (3,1) = seg 4
    # reversed, (1,2) = seg 1
    m = _read("square1")
    assert (m.npnt, m.nseg, m.ntri, m.nebd) == (4, 5, 2, 4)
    assert m.tri_seg.tolist() == [[0, 4, 3], [2, 4, 1]]
    assert m.tri_sgn.tolist() == [[1, 1, 1], [1, -1, 1]]
    assert m.ebdry_ids == [1, 2, 4, 1]
    assert m.ebdry_seg.tolist() == [0, 1, 3, 2]


@pytest.mark.parametrize("name", FILES)
def test_reference_meshes_are_consistent(name):
    m = _read(name)
    # Euler characteristic of a disc
    assert m.npnt - m.nseg + m.ntri == 1
    # the triangles tile the region the boundary encloses (shoelace over boundary sides)
    x, y = m.pnts[:, 0], m.pnts[:, 1]
    a, b = m.seg_pnt[m.ebdry_seg, 0], m.seg_pnt[m.ebdry_seg, 1]
    enclosed = 0.5 * np.sum(x[a] * y[b] - x[b] * y[a])
    assert abs(m.area2().sum() / 2 - enclosed) < 1e-12 * enclosed
    # every boundary side has exactly one triangle and is listed once
    assert len(set(m.ebdry_seg.tolist())) == m.ebdry_seg.size


@pytest.mark.parametrize("name", FILES)
def test_grd_round_trip(name, tmp_path):
    m = _read(name, essential={1: 0b011, 4: 0b011})
    p = tmp_path / "copy.grd"
    grd.write_grd(str(p), m)
    m2 = grd.read_grd(str(p), essential={1: 0b011, 4: 0b011})
    for f in ("pnts", "seg_pnt", "tri_pnt", "tri_seg", "tri_sgn", "ebdry_start", "ebdry_seg", "ebdry_essential"):
        assert np.array_equal(getattr(m, f), getattr(m2, f)), f
    assert m2.ebdry_ids == m.ebdry_ids


def test_essential_masks_by_boundary_id():
    m = _read("channel", essential={4: 0b011, 3: 0b001})
    assert m.ebdry_essential.tolist() == [0, 0, 1, 3]


def _coeffs(m, p, nv, seed):
    rng = np.random.default_rng(seed)
    im = (p - 1) * (p - 2) // 2
    return (rng.standard_normal((m.npnt, nv)), rng.standard_normal((m.nseg, p - 1, nv)),
            rng.standard_normal((m.ntri, im, nv)))


def test_restart_text_round_trip(tmp_path):
    m = _read("square4")
    v, s, i = _coeffs(m, 4, 3, 1)
    f = str(tmp_path / "rstrt_b0.txt")
    grd.write_restart_text(f, 4, v, s, i)
    text = open(f).read().splitlines()
    assert text[0] == "p0 = 4" and text[2] == "END OF HEADER"
    assert len(text) == 3 + m.npnt + 3 * m.nseg + 3 * m.ntri
    assert text[3] == "".join(f"{x:.8e}\t" for x in v[0])
    v2, s2, i2 = grd.read_restart_text(f, 4, m.npnt, m.nseg, m.ntri, 3)
    for a, b in ((v, v2), (s, s2), (i, i2)):
        assert np.abs(a - b).max() <= 5e-9 * np.abs(a).max()


def test_restart_text_order_change(tmp_path):
    """Reading a p = 4 file at p0 = 3 and p0 = 2 keeps the leading hierarchical modes
    (interior row m = 1 of p = 4 holds modes 0, 1, row m = 2 holds mode 2; p0 = 3 keeps
    mode 0 only); reading a p = 2 file at p0 = 4 zero-fills the higher modes."""
    m = _read("square1")
    v, s, i = _coeffs(m, 4, 3, 2)
    f = str(tmp_path / "p4.txt")
    grd.write_restart_text(f, 4, v, s, i)
    for p0 in (3, 2):
        v3, s3, i3 = grd.read_restart_text(f, p0, m.npnt, m.nseg, m.ntri, 3)
        assert np.allclose(v3, v, rtol=1e-8)
        assert np.allclose(s3, s[:, :p0 - 1], rtol=1e-8)
        if p0 == 3:
            assert np.allclose(i3[:, 0], i[:, 0], rtol=1e-8)
        else:
            assert i3.shape == (m.ntri, 0, 3)
    v2, s2, i2 = _coeffs(m, 2, 3, 3)
    f2 = str(tmp_path / "p2.txt")
    grd.write_restart_text(f2, 2, v2, s2, i2)
    v4, s4, i4 = grd.read_restart_text(f2, 4, m.npnt, m.nseg, m.ntri, 3)
    assert np.allclose(v4, v2, rtol=1e-8)
    assert np.allclose(s4[:, :1], s2, rtol=1e-8) and not s4[:, 1:].any()
    assert not i4.any()


def test_free_stream_residual_vanishes_on_reference_mesh(oracle_mod):
    """A uniform flow is an exact steady solution: the oracle's residual on the
    unstructured tim.1 mesh is zero to rounding (p = 4, natural boundaries)."""
    m = _read("tim.1", essential=0)
    T = oracle_mod.TriHpIns(m, log2p=2, mu=1e-2)
    v = np.tile([1.0, 0.4, 0.2], (m.npnt, 1))
    T.set_state(v, np.zeros((m.nseg, 3, 3)), np.zeros((m.ntri, 3, 3)))
    T.rsdl()
    rv, rs, ri = T.get_res()
    # the edge boundaries carry the flux through the outer boundary; the interior
    # vertices and all interior sides and elements balance exactly
    interior = ~m.boundary_vertex_mask()
    assert np.abs(rv[interior]).max() < 1e-13
    assert np.abs(ri).max() < 1e-13
    assert np.abs(rs[~m.boundary_segment_mask()]).max() < 1e-13


@pytest.mark.parametrize("edit,what", [
    (("0: 0 1", "0: 1 0"), "not stored counter-clockwise"),     # boundary side 0 reversed
    (("0: 3 0 1", "0: 3 1 0"), "counter-clockwise"),            # triangle 0 clockwise
])
def test_inconsistent_grd_is_rejected(tmp_path, edit, what):
    """A .grd whose side or triangle orientation contradicts the rest raises ValueError
    naming the invariant (the check survives python -O: no assert)."""
    src = open(os.path.join(HERE, "square1.grd")).read().splitlines()
    old, new = edit
    k = next(i for i, ln in enumerate(src) if ln.strip() == old)
    src[k] = new
    p = tmp_path / "bad.grd"
    p.write_text("\n".join(src) + "\n")
    with pytest.raises(ValueError, match=what):
        grd.read_grd(str(p))
