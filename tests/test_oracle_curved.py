"""Curved elements: the coordinate side modes of a curved boundary (hpmp.curved, the
reference's hp_edge_bdry::curv_init) and the oracle's curved branches of element_rsdl,
setup_preconditioner and the boundary-edge flux (tri_hp/ins/rsdl.cpp:83-257,
tri_hp/ins/setup_preconditioner.cpp:44-90, tri_hp/hp_boundary.cpp:973-1015).

Pins, each of which a plausible slip (a transposed metric term, a wrong sign of the
projected offset, a missing side mode on the edge) fails:
* a boundary that is a quadratic curve is reproduced exactly by the side modes;
* circular sides converge to the circle as p grows;
* with all side modes zero the curved branches give the straight branches' numbers;
* a divergence-free linear velocity field, represented exactly on the curved p = 2
  elements (its modes follow the coordinates' modes), has a continuity residual
  lf_im + edge flux that vanishes at every mode, as the divergence theorem fixes when
  the element integrals and the edge integrals see the same curved geometry (p = 2:
  the quadrature is exact for these integrands); on straight geometry the same state
  leaves a residual of order h^2;
* Taylor-Couette flow between two circles, an exact solution of the steady
  Navier-Stokes equations, is reached to 2e-6 with curved sides at p = 4 on a coarse
  ring, and missed by more than 1000 times that with straight sides.
"""
import numpy as np
import pytest

from conftest import rel_err


def _onb(r_in=0.5, r_out=1.0):
    def f(x):
        r = np.linalg.norm(x, axis=1, keepdims=True)
        return x / r * np.where(r < 0.5 * (r_in + r_out), r_in, r_out)
    return f


@pytest.mark.parametrize("p", [2, 4])
def test_side_modes_reproduce_a_quadratic_boundary(p):
    from hpmp import curved as CV, mesh as M
    m = M.unit_square(1, 0.0, 1)
    bottom = [s for s in m.ebdry_seg if np.all(m.pnts[m.seg_pnt[s], 1] == 0.0)]
    assert len(bottom) == 1
    c = 0.3

    def to_b(x):
        return np.stack([x[:, 0], c * x[:, 0] * (1.0 - x[:, 0])], axis=1)

    crds = CV.side_modes(m, bottom, to_b, p)
    pts = CV.side_points(m, bottom, crds, p)[0]
    assert np.abs(pts[:, 1] - c * pts[:, 0] * (1 - pts[:, 0])).max() < 1e-15
    # the x positions stay those of the chord (the offset is vertical)
    chord = CV.side_points(m, bottom, np.zeros_like(crds), p)[0]
    assert np.abs(pts[:, 0] - chord[:, 0]).max() < 1e-15


def test_side_modes_converge_to_the_circle():
    from hpmp import curved as CV
    m = CV.annulus(3, 16)
    bs = m.ebdry_seg
    dev = {}
    for p in (2, 4):
        crds = CV.side_modes(m, bs, _onb(), p)
        pts = CV.side_points(m, bs, crds, p).reshape(-1, 2)
        r = np.linalg.norm(pts, axis=1)
        dev[p] = np.abs(r - np.where(r < 0.75, 0.5, 1.0)).max()
    pts = CV.side_points(m, bs, np.zeros((bs.size, 1, 2)), 2).reshape(-1, 2)
    r = np.linalg.norm(pts, axis=1)
    chord = np.abs(r - np.where(r < 0.75, 0.5, 1.0)).max()
    print("max radial deviation at the Gauss points: chord", chord, "p=2", dev[2], "p=4", dev[4])
    assert dev[2] < chord / 10 and dev[4] < dev[2] / 10


@pytest.mark.parametrize("log2p", [1, 2])
def test_zero_side_modes_give_the_straight_branch(oracle_mod, log2p):
    from hpmp import curved as CV, workloads as W
    m = CV.annulus(4, 16, jitter=0.3, seed=2)
    p = 1 << log2p
    st = W.noisy_freestream(m, p, seed=3, scale=0.3)
    O1 = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    O2 = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    O2.set_curved(m.ebdry_seg, np.zeros((m.ebdry_seg.size, p - 1, 2)))
    for O in (O1, O2):
        O.set_state(*st)
        assert O.setup_preconditioner() == 0
        O.rsdl()
    assert rel_err(O2.get_res(), O1.get_res()) < 1e-14
    assert rel_err(O2.get_prcn(), O1.get_prcn()) < 1e-14


def linear_field_state(m, crds, A, p, curved_sides):
    """Modes of u = A x on the elements whose boundary sides carry `crds`: vertex values
    A x_v, side modes A crds on the curved sides (the coordinates' modes), zero
    elsewhere; pressure zero."""
    sm = p - 1
    im = (p - 1) * (p - 2) // 2
    v = np.zeros((m.npnt, 3))
    v[:, :2] = m.pnts @ A.T
    s = np.zeros((m.nseg, sm, 3))
    if curved_sides is not None:
        s[curved_sides, :, :2] = crds @ A.T
    i = np.zeros((m.ntri, im, 3))
    return v, s, i


def test_divergence_free_field_has_zero_continuity_residual_on_curved_elements(oracle_mod):
    from hpmp import curved as CV
    m = CV.annulus(4, 12, jitter=0.3, seed=5)
    p, log2p = 2, 1
    bs = m.ebdry_seg
    crds = CV.side_modes(m, bs, _onb(), p)
    A = np.array([[0.3, 1.0], [-0.7, -0.3]])
    st = linear_field_state(m, crds, A, p, bs)

    def continuity(curved):
        O = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
        if curved:
            O.set_curved(bs, crds)
        O.set_state(*st)
        O.rsdl(1)   # beta(1) = 0: lf_im + the edge flux (res_r starts at zero)
        v, s, i = O.get_res()
        return max(np.abs(v[:, 2]).max(), np.abs(s[:, :, 2]).max())

    rc, rs = continuity(True), continuity(False)
    print("continuity residual of u = A x: curved", rc, "straight sides", rs)
    assert rc < 1e-14 and rs > 1e-5


def couette_problem(m, p, crds, curved_sides):
    """Taylor-Couette flow between the circles r = 0.5 (turning at unit angular speed)
    and r = 1 (at rest): u_theta = A r + B / r, p = A^2 r^2 / 2 + 2 A B ln r - B^2 / (2 r^2).
    The walls' condition is the physical one, rigid rotation (-y, x) on the inner wall
    and rest on the outer, evaluated where the discrete sides lie: on curved sides that
    is the Couette flow's own boundary value, on straight sides the polygon turning
    inside a polygon, whose flow differs from Couette's.  Returns (exact(x) -> (u, v, p),
    the initial state: boundary vertices and the side modes of the wall data along the
    given sides, zero inside)."""
    from hpmp import curved as CV
    r0, r1 = 0.5, 1.0
    A = -r0 ** 2 / (r1 ** 2 - r0 ** 2)
    B = r0 ** 2 * r1 ** 2 / (r1 ** 2 - r0 ** 2)

    def exact(x):
        r = np.linalg.norm(x, axis=1)
        ut = A * r + B / r
        pr = 0.5 * A ** 2 * r ** 2 + 2 * A * B * np.log(r) - 0.5 * B ** 2 / r ** 2
        return np.stack([-ut * x[:, 1] / r, ut * x[:, 0] / r, pr], axis=1)

    sm, im = p - 1, (p - 1) * (p - 2) // 2
    v = np.zeros((m.npnt, 3)); s = np.zeros((m.nseg, sm, 3)); i = np.zeros((m.ntri, im, 3))
    bv = m.boundary_vertex_mask()
    v[bv, :2] = exact(m.pnts[bv])[:, :2]
    bs = m.ebdry_seg
    def wall(x):
        inner = np.linalg.norm(x, axis=1) < 0.5 * (r0 + r1)
        return np.where(inner[:, None], np.stack([-x[:, 1], x[:, 0]], axis=1), 0.0)

    s[bs, :, :2] = CV.side_data_modes(m, bs, crds if curved_sides else np.zeros_like(crds), wall, p)
    return exact, (v, s, i)


def test_taylor_couette_flow_with_curved_sides(oracle_mod):
    from hpmp import curved as CV
    m = CV.annulus(4, 24)
    p = 4
    crds = CV.side_modes(m, m.ebdry_seg, _onb(), p)
    bv, bs = m.boundary_vertex_mask(), m.boundary_segment_mask()

    def judge(O):
        O.rsdl()
        rv, rs, ri = O.get_res()
        rv[bv, :2] = 0.0
        rs[bs, :, :2] = 0.0
        return max(np.abs(rv).max(), np.abs(rs).max(), np.abs(ri).max())

    err = {}
    for curved in (True, False):
        exact, st = couette_problem(m, p, crds, curved)
        O = oracle_mod.TriHpIns(m, log2p=2, mu=0.1)
        if curved:
            O.set_curved(m.ebdry_seg, crds)
        O.set_state(*st)
        r0 = judge(O)
        for it in range(3000):
            O.cycle(vw=1, mglvls=3, itercrsn=1, iterrfne=1)
            if it % 20 == 19 and judge(O) < 1e-11 * r0:
                break
        assert judge(O) < 1e-10 * r0
        v = O.get_state()[0]
        err[curved] = np.abs(v[~bv, :2] - exact(m.pnts[~bv])[:, :2]).max()
    print("vertex velocity error: curved", err[True], "straight", err[False])
    assert err[True] < 2e-6 and 1000 * err[True] < err[False]
