"""Pins of the tet_basis / tet_hp_ins oracle (p = 1, 2) by mathematics.

Basis: the Gauss rules are the Legendre / Jacobi(1,0) / Jacobi(2,0) rules of numpy and
scipy scaled as tet_basis/initialize.cpp:136-295 scales them; the vertex modes are
the barycentric coordinates of the reference tet (V2 at r=s=t=-1, V3 at r=1, V1 at
s=1, V0 at t=1); the p=2 edge modes are the products of their end vertices'
barycentrics for the local edge table of tet_mesh/setup.cpp:230-235; proj's r, s, t
derivatives, derivr/s/t and intgrtrst agree with those closed forms; the mass matrix
equals the closed-form integrals of barycentric monomials; vdiag and sfmv satisfy
the defining mass-orthogonality of tet_basis/initialize.cpp:642-800.
Residual: a uniform stream leaves no interior residual and no continuity residual
anywhere once the inflow face flux is added; the continuity residual sums to zero
for any state (conservation); the relaxation converges and p-multigrid reaches the
same solution.
"""
from math import factorial

import numpy as np
import pytest
from scipy.special import roots_jacobi

from hpmp import tetmesh as TM

VS = [(2, 3), (3, 1), (2, 1), (1, 0), (2, 0), (3, 0)]


@pytest.fixture(scope="module")
def OT(oracle_mod):
    import oracle_tet
    return oracle_tet


def _rst(B):
    xp, yp, zp = B.table("xp"), B.table("yp"), B.table("zp")
    X, Y, Z = np.meshgrid(xp, yp, zp, indexing="ij")
    t = Z.ravel()
    s = (1 + Y.ravel()) * (1 - t) / 2 - 1
    r = (1 + X.ravel()) * (-s - t) / 2 - 1
    return r, s, t


def _bary(r, s, t):
    l0, l1, l3 = (1 + t) / 2, (1 + s) / 2, (1 + r) / 2
    return np.stack([l0, l1, 1 - l0 - l1 - l3, l3])


GRAD = np.array([[0, 0, .5], [0, .5, 0], [-.5, -.5, -.5], [.5, 0, 0]])  # d lambda_v / d(r,s,t)


def _modes(B, r, s, t):
    lam = _bary(r, s, t)
    val = [lam[v] for v in range(4)]
    grad = [np.tile(GRAD[v][:, None], (1, r.size)) for v in range(4)]
    if B.p == 2:
        for a, b in VS:
            val.append(lam[a] * lam[b])
            grad.append(lam[a] * GRAD[b][:, None] + lam[b] * GRAD[a][:, None])
    return np.array(val), np.array(grad)


@pytest.mark.parametrize("p", [1, 2])
def test_gauss_rules(OT, p):
    B = OT.TetBasis(p)
    n = p + 1
    x, w = np.polynomial.legendre.leggauss(n)
    assert np.allclose(B.table("xp"), x, atol=1e-15) and np.allclose(B.table("wtx"), w, atol=1e-15)
    y, wy = roots_jacobi(n, 1.0, 0.0)
    assert np.allclose(B.table("yp"), y, atol=1e-14) and np.allclose(B.table("wty"), 0.5 * wy, atol=1e-14)
    z, wz = roots_jacobi(n, 2.0, 0.0)
    assert np.allclose(B.table("zp"), z, atol=1e-14) and np.allclose(B.table("wtz"), 0.25 * wz, atol=1e-14)


@pytest.mark.parametrize("p", [1, 2])
def test_modes_and_derivatives_are_the_closed_forms(OT, p):
    B = OT.TetBasis(p)
    r, s, t = _rst(B)
    val, grad = _modes(B, r, s, t)
    for m in range(B.tm):
        f, dx, dy, dz = B.proj(np.eye(B.tm)[m])
        assert np.abs(f - val[m]).max() < 1e-14
        assert np.abs(dx - grad[m][0]).max() < 1e-13
        assert np.abs(dy - grad[m][1]).max() < 1e-13
        assert np.abs(dz - grad[m][2]).max() < 1e-13
        assert np.abs(B.proj_val(np.eye(B.tm)[m]) - f).max() == 0.0
    u = [0.3, -1.2, 0.7, 2.0]
    assert np.abs(B.proj_lin(*u) - sum(u[v] * val[v] for v in range(4))).max() < 1e-14


def test_gauss_point_derivatives_are_exact_on_quadratics(OT):
    B = OT.TetBasis(2)
    r, s, t = _rst(B)
    f = 0.3 + r - 2 * s + 0.5 * t + r * s - 0.7 * t * t + 0.4 * r * t + s * s
    fr, fs, ft = 1 + s + 0.4 * t, -2 + r + 2 * s, 0.5 - 1.4 * t + 0.4 * r
    assert np.abs(B.deriv("r", f) - fr).max() < 1e-12
    assert np.abs(B.deriv("s", f) - fs).max() < 1e-12
    assert np.abs(B.deriv("t", f) - ft).max() < 1e-12
    # accumulation into the output, as the reference's loops do
    assert np.abs(B.deriv("r", f, init=np.ones(B.G)) - (fr + 1)).max() < 1e-12


def _tet_monomial(a, b, c, d):
    """Integral over the reference tet (volume 4/3) of l0^a l1^b l2^c l3^d."""
    return 6 * factorial(a) * factorial(b) * factorial(c) * factorial(d) / factorial(3 + a + b + c + d) * (4 / 3)


@pytest.mark.parametrize("p", [1, 2])
def test_mass_matrix_and_integration(OT, p):
    B = OT.TetBasis(p)
    mm = B.table("mm").reshape(B.tm, B.tm)
    ends = [(v,) for v in range(4)] + (VS if p == 2 else [])
    for i, ei in enumerate(ends):
        for j, ej in enumerate(ends):
            pw = [0, 0, 0, 0]
            for v in ei + ej:
                pw[v] += 1
            assert abs(mm[i, j] - _tet_monomial(*pw)) < 1e-14
    # intgrtrst is minus the weighted gradient pairing
    r, s, t = _rst(B)
    _, grad = _modes(B, r, s, t)
    rng = np.random.default_rng(3)
    dx, dy, dz = rng.standard_normal((3, B.G))
    w = np.einsum("i,j,k->ijk", B.table("wtx"), B.table("wty"), B.table("wtz")).ravel()
    # collapse Jacobian is in the Jacobi weights: sum w = 4/3
    assert abs(w.sum() - 4 / 3) < 1e-14
    exp = -np.array([(w * (g[0] * dx + g[1] * dy + g[2] * dz)).sum() for g in grad])
    assert np.abs(B.intgrtrst(dx, dy, dz) - exp).max() < 1e-13
    f = rng.standard_normal(B.G)
    val, _ = _modes(B, r, s, t)
    assert np.abs(B.intgrt(f) - (w * val * f).sum(axis=1)).max() < 1e-14


def test_face_transforms(OT):
    B = OT.TetBasis(2)
    x, y = B.table("xp"), B.table("yp")
    X, Y = np.meshgrid(x, y, indexing="ij")
    s = Y.ravel()
    r = (1 + X.ravel()) * (1 - s) / 2 - 1
    l1, l3 = (1 + s) / 2, (1 + r) / 2   # face vertices V1 (s=1), V3 (r=1), V2
    lam = [l1, 1 - l1 - l3, l3]
    vals = lam + [lam[1] * lam[2], lam[0] * lam[2], lam[0] * lam[1]]   # E1 (V2,V3), E2 (V1,V3), E3 (V1,V2)
    for m in range(6):
        f, _, _ = B.proj2d(np.eye(6)[m])
        assert np.abs(f - vals[m]).max() < 1e-14
    w = np.einsum("i,j->ij", B.table("wtx"), B.table("wty")).ravel()
    g = np.random.default_rng(5).standard_normal(w.size)
    assert np.abs(B.intgrt2d(g) - np.array([(w * v * g).sum() for v in vals])).max() < 1e-14


def test_lumped_vertex_factors(OT):
    B = OT.TetBasis(2)
    mm = B.table("mm").reshape(10, 10)
    sf = B.table("sfmv")
    assert abs(sf[0] - sf[1]) < 1e-14
    # vertex 3 plus -sfmv times its three edges is mass-orthogonal to vertices 0, 1, 2
    comb = np.zeros(10)
    comb[3] = 1.0
    comb[[4, 5, 9]] = -sf[0]
    assert np.abs(mm[:3] @ comb).max() < 1e-14
    # vdiag: inverse of vertex 0's self-mass after the same removal (edges 7, 8, 9)
    c0 = np.zeros(10)
    c0[0] = 1.0
    assert abs(1.0 / B.table("vdiag")[0] - (mm[0, 0] - sf[0] * (mm[7, 0] + mm[8, 0] + mm[9, 0]))) < 1e-14
    assert B.table("ediag")[0] == 100.0


def _state(m, em0, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    v = np.zeros((m.npnt, 4))
    v[:, 0], v[:, 1], v[:, 2], v[:, 3] = 1.0, 0.3, -0.2, 0.5
    e = np.zeros((m.nseg, em0, 4))
    v2 = v + amp * rng.standard_normal(v.shape)
    e2 = amp * rng.standard_normal(e.shape)
    v2[m.boundary_vertex_mask(), :3] = v[m.boundary_vertex_mask(), :3]
    e2[m.boundary_edge_mask(), :, :3] = 0.0
    return (v, e), (v2, e2)


@pytest.mark.parametrize("log2p", [0, 1])
def test_uniform_stream_and_conservation(OT, log2p):
    m = TM.unit_cube(3, 0.2, 1)
    TM.check_mesh(m)
    O = OT.TetHpIns(m, log2p=log2p, mu=1e-2)
    (v, e), (v2, e2) = _state(m, (1 << log2p) - 1, 1)
    O.set_state(v, e)
    assert O.setup_preconditioner() == 0
    O.rsdl()
    rv, re = O.get_res()
    bv, be = m.boundary_vertex_mask(), m.boundary_edge_mask()
    assert np.abs(rv[~bv]).max() < 1e-14
    if re.size:
        assert np.abs(re[~be]).max() < 1e-14
    assert np.abs(rv[:, 3]).max() < 1e-14 and (not re.size or np.abs(re[:, :, 3]).max() < 1e-14)
    O.set_state(v2, e2)
    O.setup_preconditioner()
    for stage in (0, 2, 5):
        O.rsdl(stage)
        rv, _ = O.get_res()
        assert abs(rv[:, 3].sum()) < 1e-13


def test_relaxation_and_p_multigrid_converge_to_one_solution(OT):
    m = TM.unit_cube(3, 0.2, 1)
    _, (v2, e2) = _state(m, 1, 2)
    sols = []
    for mg in (1, 2):
        O = OT.TetHpIns(m, log2p=1, mu=5e-2)
        O.set_state(v2, e2)
        for it in range(1500):
            O.setup_preconditioner()
            O.update()
            if mg == 2:
                O.rsdl()
                O.mg_restrict()
                O.setup_preconditioner()
                O.update()
                O.mg_prolongate()
            if it % 100 == 99:
                O.rsdl()
                rv, re = O.get_res()
                rv[m.boundary_vertex_mask(), :3] = 0
                re[m.boundary_edge_mask(), :, :3] = 0
                if max(np.abs(rv).max(), np.abs(re).max()) < 1e-13:
                    break
        sols.append((it, O.get_state()))
    print("iterations: relaxation", sols[0][0] + 1, "p-multigrid", sols[1][0] + 1)
    assert sols[1][0] < sols[0][0]
    a, b = sols
    # a constant pressure is a null mode of velocity-Dirichlet problems
    dv = a[1][0] - b[1][0]
    dv[:, 3] -= dv[:, 3].mean()
    de = a[1][1] - b[1][1]
    assert max(np.abs(dv).max(), np.abs(de).max()) < 1e-10


def test_nan_state_reports_setup_error_in_max_order(OT, oracle_mod):
    """setup_preconditioner returns 1 when qmax is NaN (the reference's
    `!(qmax >= 0.0)` check, tet_hp/ins/setup_preconditioner.cpp:59-76).  qmax is
    accumulated with MAX of sys/param.h, ((a) > (b) ? (a) : (b)), so a NaN q survives
    only as the last vertex of an element: a NaN velocity at a vertex is caught exactly
    when that vertex is the last one of some element.  std::max, or the operands the
    other way round, would change which vertices are caught."""
    from hpmp import mesh as M, workloads as W
    m = TM.unit_cube(2)
    v = np.tile([1.0, 0.0, 0.0, 0.0], (m.npnt, 1))
    e = np.zeros((m.nseg, 1, 4))
    O = OT.TetHpIns(m, log2p=1, mu=1e-2)
    O.set_state(v, e)
    assert O.setup_preconditioner() == 0
    last = set(m.tet_pnt[:, 3].tolist())
    assert 0 < len(last) < m.npnt          # both outcomes occur on this mesh
    for pos in range(m.npnt):
        w = v.copy()
        w[pos, pos % 3] = np.nan
        O.set_state(w, e)
        assert O.setup_preconditioner() == (1 if pos in last else 0), pos
    t = M.unit_square(3, 0.1, 1)
    st = list(W.noisy_freestream(t, 4, seed=1))
    Ot = oracle_mod.TriHpIns(t, log2p=2, mu=1e-2)
    Ot.set_state(*st)
    assert Ot.setup_preconditioner() == 0
    # tri_hp_ins takes qmax over the Gauss points of the projected velocity
    # (tri_hp/ins/setup_preconditioner.cpp:98-125): every point sees every vertex
    # mode, so a NaN at any vertex reaches the last point and is caught
    for pos in range(t.npnt):
        w = [x.copy() for x in st]
        w[0][pos, 1] = np.nan
        Ot.set_state(*w)
        assert Ot.setup_preconditioner() == 1, pos


@pytest.mark.parametrize("cns", [False, True])
def test_tet_element_jacobian_matches_central_differences(cns):
    """The forward-difference tet element Jacobian (tet_hp/update.cpp:134-178, one dw for
    every variable) against central differences of element_rsdl built here from the
    state, and its dw convention: with eps_r = 0 every column uses dw = eps_a."""
    import oracle_tet as OT
    from hpmp import tetmesh as TM
    m = TM.unit_cube(2, 0.15, 3)
    rng = np.random.default_rng(12)
    if cns:
        st = np.array([1.0, 0.3, 0.1, -0.05, 1.0])
        X = OT.TetHpCns(m, log2p=1, mu=1e-2, bc=(0.0, *st[1:]), gamma=1.4, prandtl=0.72, R=1.0)
    else:
        st = np.array([1.0, 0.3, -0.2, 0.5])
        X = OT.TetHpIns(m, log2p=1, mu=1e-2)
    nv = st.size
    v = np.tile(st, (m.npnt, 1)) + 0.02 * rng.standard_normal((m.npnt, nv))
    e = 0.02 * rng.standard_normal((m.nseg, 1, nv))
    X.set_state(v, e)
    assert X.setup_preconditioner() == 0
    t = 5
    K = X.element_jacobian(t)
    uht = np.zeros((nv, 10))
    uht[:, :4] = v[m.tet_pnt[t]].T
    uht[:, 4:] = e[m.tet_seg[t], 0, :].T   # edge mode 0 carries sgn^0 = 1
    h = 1e-6
    Kc = np.zeros_like(K)
    for mode in range(10):
        for var in range(nv):
            up = uht.copy(); up[var, mode] += h
            dn = uht.copy(); dn[var, mode] -= h
            rp = sum(X.element_rsdl(t, 0, up)); rm = sum(X.element_rsdl(t, 0, dn))
            Kc[:, mode * nv + var] = ((rp - rm) / (2 * h)).T.ravel()
    assert np.abs(K - Kc).max() < 1e-4 * np.abs(Kc).max()
    # dw = eps_r * sum |vertex coefficients| + eps_a, the same for all variables: a
    # column formed with eps_r = 0 and eps_a = dw matches the default one exactly
    dw = 1e-6 * np.abs(uht[:, :4]).T.sum() + 1e-10
    K2 = X.element_jacobian(t, 0.0, dw)
    assert np.abs(K2 - K).max() <= 1e-9 * np.abs(K).max()
