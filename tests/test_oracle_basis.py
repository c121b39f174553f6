"""Pins of the oracle's tri_basis restatement against mathematics that does not
go through its own formulas: library Gauss rules, scipy's Jacobi polynomials,
brute-force quadrature on the triangle, and the reference's own self-checks
(tri_basis/initialize.cpp:34-156)."""
import numpy as np
import pytest
from scipy.special import eval_jacobi, roots_jacobi, jacobi

PS = (1, 2, 4)


@pytest.mark.parametrize("p", PS)
def test_gauss_rules_match_library(oracle_mod, p):
    B = oracle_mod.Basis(p)
    xl, wl = np.polynomial.legendre.leggauss(B.gpx)
    assert np.abs(B.table("xp") - xl).max() < 1e-15
    assert np.abs(B.table("wtx") - wl).max() < 2e-15
    xj, wj = roots_jacobi(B.gpn, 1.0, 0.0)
    assert np.abs(B.table("np") - xj).max() < 1e-15
    assert np.abs(B.table("wtn") - 0.5 * wj).max() < 2e-15


@pytest.mark.parametrize("p", PS)
def test_gauss_rules_exactness(oracle_mod, p):
    B = oracle_mod.Basis(p)
    x, w = B.table("xp"), B.table("wtx")
    e, we = B.table("np"), B.table("wtn")
    for k in range(2 * B.gpx):
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs((w * x ** k).sum() - exact) < 1e-14
    for k in range(2 * B.gpn):
        # 0.5 * int_{-1}^{1} (1 - eta) eta^k
        ex = 0.5 * ((1 - (-1) ** (k + 1)) / (k + 1) - (1 - (-1) ** (k + 2)) / (k + 2))
        assert abs((we * e ** k).sum() - ex) < 1e-14


@pytest.mark.parametrize("al,be", [(0.0, 0.0), (2.0, 2.0), (1.0, 0.0), (5.0, 2.0), (7.0, 2.0)])
def test_monic_recurrence_against_scipy(oracle_mod, al, be):
    n = 6
    a, b = oracle_mod.recur(n + 1, 6 if (al, be) != (0.0, 0.0) else 1, al, be)
    xs = np.linspace(-0.97, 0.93, 11)
    pkm, pk = np.zeros_like(xs), np.ones_like(xs)
    for k in range(n):
        lead = jacobi(k, al, be).coeffs[0]
        assert np.allclose(pk, eval_jacobi(k, al, be, xs) / lead, rtol=1e-12, atol=1e-13)
        pkm, pk = pk, (xs - a[k]) * pk - (b[k] * pkm if k else 0.0)
    # b[0] is the total mass of (1-x)^al (1+x)^be on [-1, 1]
    xq, wq = roots_jacobi(20, al, be)
    assert abs(b[0] - wq.sum()) < 1e-13 * b[0]


def _basis_closed_form(p, x, eta):
    """Mode values from the basis definition (tri_basis/ptvalues.cpp:14-81) written
    with scipy's Jacobi polynomials and normalised by brute-force L2 norms."""
    sm = p - 1
    xs, ws = np.polynomial.legendre.leggauss(p + 1)

    def side1d(k, t):
        P = eval_jacobi(k, 2, 2, t)
        return 0.25 * (1 - t) * (1 + t) * P

    nrm = [1.0 / np.sqrt((ws * side1d(k, xs) ** 2).sum()) for k in range(sm)]
    gx = [np.ones_like(x), 0.5 * (1 - x), 0.5 * (1 + x)] + [side1d(k, x) * nrm[k] for k in range(sm)]
    out = [gx[0] * 0.5 * (1 + eta), gx[1] * 0.5 * (1 - eta), gx[2] * 0.5 * (1 - eta)]
    out += [gx[3 + k] * (0.5 * (1 - eta)) ** (k + 2) for k in range(sm)]
    out += [gx[2] * side1d(k, eta) * nrm[k] for k in range(sm)]
    out += [gx[1] * (-1) ** k * side1d(k, eta) * nrm[k] for k in range(sm)]
    ej, wj = roots_jacobi(p + 1, 1.0, 0.0)
    for g in range(2, sm + 1):
        for k in range(sm + 1 - g):
            f = lambda t: (0.5 * (1 - t)) ** g * 0.5 * (1 + t) * eval_jacobi(k, 2 * g + 1, 2, t)
            inrm = 1.0 / np.sqrt((0.5 * wj * f(ej) ** 2).sum())
            out.append(gx[g + 1] * f(eta) * inrm)
    return np.array(out)


@pytest.mark.parametrize("p", PS)
def test_basis_values_closed_form(oracle_mod, p):
    B = oracle_mod.Basis(p)
    rng = np.random.default_rng(3)
    for _ in range(8):
        x, eta = rng.uniform(-0.95, 0.95, 2)
        gx, _, gn, _ = B.ptvalues_deriv(x, eta)
        # oracle value of mode m = gx(xmode(m)) * gn(m)
        sm = p - 1
        xm = [0, 1, 2] + [3 + k for k in range(sm)] + [2] * sm + [1] * sm
        for g in range(2, sm + 1):
            xm += [g + 1] * (sm + 1 - g)
        vals = np.array([gx[xm[m]] * gn[m] for m in range(B.tm)])
        assert np.abs(vals - _basis_closed_form(p, np.array(x), np.array(eta))).max() < 1e-13


@pytest.mark.parametrize("p", PS)
def test_proj_derivatives_match_interpolant_derivatives(oracle_mod, p):
    """The reference's T1 self-test: proj's r/s derivatives equal derivr/derivs of the values."""
    B = oracle_mod.Basis(p)
    for m in range(B.tm):
        u = np.zeros(B.tm); u[m] = 1.0
        f, dx, dy = B.proj(u)
        assert np.abs(B.derivr(f) - dx).max() < 1e-12 * max(1.0, np.abs(dx).max())
        assert np.abs(B.derivs(f) - dy).max() < 1e-12 * max(1.0, np.abs(dy).max())


def _tri_rule(n):
    """Collapsed-coordinate rule on the reference triangle (-1,-1),(1,-1),(-1,1), exact to high degree."""
    x, wx = np.polynomial.legendre.leggauss(n)
    e, we = roots_jacobi(n, 1.0, 0.0)
    X, E = np.meshgrid(x, e, indexing="ij")
    W = np.outer(wx, 0.5 * we)
    r = 0.5 * (1 + X) * (1 - E) - 1
    return X.ravel(), E.ravel(), r.ravel(), E.ravel(), W.ravel()


@pytest.mark.parametrize("p", PS)
def test_mass_matrix_brute_force(oracle_mod, p):
    B = oracle_mod.Basis(p)
    X, E, _, _, W = _tri_rule(12)
    phi = _basis_closed_form(p, X, E)
    Mbf = (phi * W) @ phi.T
    for m in range(B.tm):
        u = np.zeros(B.tm); u[m] = 1.0
        assert np.abs(B.intgrt(B.proj_val(u)) - Mbf[m]).max() < 1e-14 * 50
    assert abs(W.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("p", PS)
def test_intgrtrs_brute_force(oracle_mod, p):
    """intgrtrs(dx, dy) = -sum_g w_g (dphi/dr dx + dphi/ds dy), derivatives by central
    differences of the closed-form basis in (r, s)."""
    B = oracle_mod.Basis(p)
    xg, eg = B.table("xp"), B.table("np")
    wg = np.outer(B.table("wtx"), B.table("wtn")).ravel()
    X, E = np.meshgrid(xg, eg, indexing="ij")
    R = 0.5 * (1 + X) * (1 - E) - 1
    S = E
    h = 1e-6

    def phi_rs(r, s):
        x = 2 * (1 + r) / (1 - s) - 1
        return _basis_closed_form(p, x, s)

    dr = (phi_rs(R + h, S) - phi_rs(R - h, S)) / (2 * h)
    ds = (phi_rs(R, S + h) - phi_rs(R, S - h)) / (2 * h)
    rng = np.random.default_rng(5)
    fx = rng.standard_normal((B.gpx, B.gpn)); fy = rng.standard_normal((B.gpx, B.gpn))
    ref = -((dr.reshape(B.tm, -1) * fx.ravel() + ds.reshape(B.tm, -1) * fy.ravel()) * wg).sum(axis=1)
    assert np.abs(B.intgrtrs(fx, fy) - ref).max() < 1e-7


@pytest.mark.parametrize("p", (2, 4))
def test_lumped_mass_constraints(oracle_mod, p):
    """The vertex/side functions built from sfmv/ifmb are M-orthogonal to the modes the
    construction constrains (tri_basis/initialize.cpp:640-760), with M by brute force."""
    B = oracle_mod.Basis(p)
    sm, im, bm, tm = B.sm, B.im, B.bm, B.tm
    X, E, _, _, W = _tri_rule(12)
    phi = _basis_closed_form(p, X, E)
    M = (phi * W) @ phi.T
    sfmv = B.table("sfmv").reshape(2, sm)
    ifmb = B.table("ifmb").reshape(bm, im) if im else np.zeros((bm, 0))
    # vertex 2 function: sides 0 (i1 = 0) and 1 (i2 = 1) corrected, interiors by ifmb(2,:)
    u = np.zeros(tm); u[2] = 1.0
    u[3:3 + sm] = -sfmv[1]
    u[3 + sm:3 + 2 * sm] = -sfmv[0]
    u[bm:] = -ifmb[2]
    Mu = M @ u
    assert abs(Mu[0]) < 1e-13 and abs(Mu[1]) < 1e-13
    for k in range(3):
        assert np.abs(Mu[3 + k * sm:3 + k * sm + sm - 1]).max(initial=0) < 1e-13
    assert abs(Mu[2] - float(B.table("vdiag")[0])) < 1e-13
    if im:
        bfmi = B.table("bfmi").reshape(bm, im)
        Mii = M[bm:, bm:]
        assert np.abs(Mii @ bfmi.T - M[bm:, :bm]).max() < 1e-13


def test_dpbtrsnu2_literal_indexing(oracle_mod):
    """utilities/myblas.c:91-122 uses b[j] (not b[i-1]) once i >= ofdg; on a diagonal
    band factor it is the exact solve."""
    L = oracle_mod.lib()
    ab = np.array([[0.0, 2.0], [0.0, 3.0], [0.0, 4.0]])
    bb = np.array([2.0, 9.0, 32.0])
    L.oracle_dpbtrsnu2(ab.ravel().copy(), 2, 3, 1, bb, 1)
    assert np.allclose(bb, [0.5, 1.0, 2.0])
    # with a super-diagonal, row 2 couples to b[0] (the literal indexing)
    ab2 = np.array([[0.0, 1.0], [0.5, 1.0], [0.5, 1.0]])
    bb = np.array([1.0, 1.0, 1.0])
    L.oracle_dpbtrsnu2(ab2.ravel().copy(), 2, 3, 1, bb, 1)
    f = np.array([1.0, 1.0 - 0.5 * 1.0, 1.0 - 0.5 * 1.0])      # forward with b[0]
    x = f.copy()
    x[2] = f[2]; x[0] -= 0.5 * x[2]
    x[1] = f[1]; x[0] -= 0.5 * x[1]
    assert np.allclose(bb, x)
