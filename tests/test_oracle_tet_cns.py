"""Pins of the tet_hp_cns oracle (oracle/otet_cns.cpp) by mathematics that does not
restate its formulas: free-stream preservation, conservation, rotation covariance,
the inviscid weak form against an independent quadrature, the flux-Jacobian algebra
of the Weiss-Smith tau against finite-difference Jacobians, and convergence."""
import numpy as np
import pytest

from hpmp import tetmesh as TM

GAM = 1.4
VS = [(2, 3), (3, 1), (2, 1), (1, 0), (2, 0), (3, 0)]
STATE = np.array([1.0, 0.3, 0.1, -0.05, 1.0])   # p, u, v, w, RT


@pytest.fixture(scope="module")
def OT(oracle_mod):
    import oracle_tet
    return oracle_tet


def _cons(w):
    p, u, v, ww, rt = w
    rho = p / rt
    return np.array([rho, rho * u, rho * v, rho * ww, rho * (rt / (GAM - 1) + 0.5 * (u * u + v * v + ww * ww))])


def _flux(w, k):
    p, u, v, ww, rt = w
    rho, vel = p / rt, np.array([u, v, ww])
    h = GAM / (GAM - 1) * rt + 0.5 * vel @ vel
    f = np.zeros(5)
    f[0] = rho * vel[k]
    for i in range(3):
        f[1 + i] = rho * vel[k] * vel[i] + (p if i == k else 0.0)
    f[4] = rho * vel[k] * h
    return f


def _jac(fun, w, eps=1e-6):
    J = np.zeros((5, 5))
    for j in range(5):
        d = np.zeros(5)
        d[j] = eps * max(1.0, abs(w[j]))
        J[:, j] = (fun(w + d) - fun(w - d)) / (2 * d[j])
    return J


def _mk(OT, m, log2p=1, mu=1e-2, bc=STATE, **kw):
    return OT.TetHpCns(m, log2p=log2p, mu=mu, gamma=GAM, prandtl=0.72, R=1.0, bc=(0.0, *bc[1:]), **kw)


def _random_state(m, em0, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    v = np.tile(STATE, (m.npnt, 1)) + amp * rng.standard_normal((m.npnt, 5))
    e = amp * rng.standard_normal((m.nseg, em0, 5))
    return v, e


@pytest.mark.parametrize("log2p", [0, 1])
@pytest.mark.parametrize("mu", [0.0, 2e-2])
def test_free_stream_is_preserved(OT, log2p, mu):
    """A uniform state equal to the inflow values: element and face fluxes cancel in
    every equation, every mode and every stage."""
    m = TM.unit_cube(3, 0.2, 1)
    st = np.array([1.0e5, 30.0, -10.0, 5.0, 287.058 * 300.0])
    O = OT.TetHpCns(m, log2p=log2p, mu=mu, bc=tuple(st))
    O.set_state(np.tile(st, (m.npnt, 1)), np.zeros((m.nseg, (1 << log2p) - 1, 5)))
    assert O.setup_preconditioner() == 0
    scale = np.array([1e5 * 0.1, 1e5 * 0.1, 1e5 * 0.1, 1e5 * 0.1, 1e6])   # flux magnitudes x face area
    for stage in (0, 1, 5):
        O.rsdl(stage)
        rv, re = O.get_res()
        assert np.all(np.abs(rv).max(axis=0) < 1e-13 * scale)
        if re.size:
            assert np.all(np.abs(re).max(axis=(0, 1)) < 1e-13 * scale)


@pytest.mark.parametrize("log2p", [0, 1])
def test_conservation(OT, log2p):
    """The vertex modes form a partition of unity, so the element terms of every
    conservation law sum to zero over the vertices: sum_v res(v, n) = the net boundary
    flux of equation n through the inflow faces."""
    m = TM.unit_cube(3, 0.2, 2)
    O = _mk(OT, m, log2p)
    O.set_state(*_random_state(m, (1 << log2p) - 1, 4))
    assert O.setup_preconditioner() == 0
    face = sum(O.face_rsdl(f)[:, :3].sum(axis=1) for f in range(m.nbf))
    for stage in (5, 0, 1):
        O.rsdl(stage)
        rv, _ = O.get_res()
        tot = rv.sum(axis=0)
        assert np.all(np.abs(tot - face) < 1e-12 * np.abs(rv).max(axis=0)), (stage, tot - face)


def _rotation():
    a, b, c = 0.7, -0.4, 1.1
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rx = np.array([[1, 0, 0], [0, np.cos(c), -np.sin(c)], [0, np.sin(c), np.cos(c)]])
    return Rz @ Ry @ Rx


@pytest.mark.parametrize("log2p", [0, 1])
def test_rotation_covariance(OT, log2p):
    """Rotating mesh, velocities and inflow velocity rotates the momentum residual and
    leaves the mass and energy residuals unchanged.  tau = alpha I keeps the
    stabilisation covariant (the Weiss-Smith tau is built from per-axis maxima)."""
    m = TM.unit_cube(2, 0.2, 3)
    Rm = _rotation()
    mr = TM.from_tets(m.pnts @ Rm.T, m.tet_pnt)
    assert np.array_equal(mr.tet_pnt, m.tet_pnt) and np.array_equal(mr.tet_seg, m.tet_seg)
    em0 = (1 << log2p) - 1
    v, e = _random_state(m, em0, 5)
    bc = STATE.copy()
    vr, er, bcr = v.copy(), e.copy(), bc.copy()
    vr[:, 1:4] = v[:, 1:4] @ Rm.T
    if em0:
        er[:, :, 1:4] = e[:, :, 1:4] @ Rm.T
    bcr[1:4] = Rm @ bc[1:4]
    res = []
    for mesh, st, b in ((m, (v, e), bc), (mr, (vr, er), bcr)):
        O = _mk(OT, mesh, log2p, bc=b)
        O.set_state(*st)
        O.setup_preconditioner()
        O.set_tau(np.tile(0.3 * np.eye(5), (mesh.ntet, 1, 1)))
        out = []
        for stage in (0, 1):
            O.rsdl(stage)
            out.append(O.get_res())
        res.append(out)
    for k in range(2):
        (a, ae), (b, be) = res[0][k], res[1][k]
        scale = np.abs(a).max()
        assert np.abs(b[:, [0, 4]] - a[:, [0, 4]]).max() < 1e-12 * scale
        assert np.abs(b[:, 1:4] - a[:, 1:4] @ Rm.T).max() < 1e-12 * scale
        if em0:
            assert np.abs(be[:, :, 1:4] - ae[:, :, 1:4] @ Rm.T).max() < 1e-12 * scale
            assert np.abs(be[:, :, [0, 4]] - ae[:, :, [0, 4]]).max() < 1e-12 * scale


def test_inviscid_weak_form_against_an_independent_quadrature(OT):
    """lf_im(n, i) = -int_T grad(phi_i) . F_n(u) dV for a linear state with uniform RT on
    the p = 2 basis (integrand of degree <= 5, which the 27-point collapsed rule
    integrates exactly),
    against a 10^3-point Duffy-mapped Gauss-Legendre rule on the physical tet; phi are
    the barycentric vertex modes and the products lambda_a lambda_b on the edges."""
    m = TM.unit_cube(1, 0.0, 0)
    O = _mk(OT, m, 1, mu=0.0)
    rng = np.random.default_rng(7)
    t = 2
    X = m.pnts[m.tet_pnt[t]]                       # 4 x 3
    wv = np.tile(STATE, (4, 1)) + 0.2 * rng.standard_normal((4, 5))
    wv[:, 0] = 1.0 + 0.3 * rng.random(4)
    wv[:, 4] = 1.1                                 # uniform RT keeps rho = p / RT linear
    uht = np.zeros((5, 10))
    uht[:, :4] = wv.T
    _, lim = O.element_rsdl(t, 1, uht)
    # barycentric gradients of the affine tet
    Tm = np.vstack([X.T, np.ones(4)])              # [x; 1] = Tm lambda
    Ti = np.linalg.inv(Tm)
    glam = Ti[:, :3]                               # d lambda_v / dx
    vol = abs(np.linalg.det(np.vstack([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6.0)
    xg, wg = np.polynomial.legendre.leggauss(10)
    xg, wg = 0.5 * (xg + 1), 0.5 * wg
    ref = np.zeros((5, 10))
    for a, wa in zip(xg, wg):
        for b, wb in zip(xg, wg):
            for c, wc in zip(xg, wg):
                l1, l2, l3 = a, b * (1 - a), c * (1 - a) * (1 - b)
                lam = np.array([1 - l1 - l2 - l3, l1, l2, l3])     # barycentrics of the tet's vertices
                jw = wa * wb * wc * (1 - a) ** 2 * (1 - b) * 6.0 * vol
                w = lam @ wv
                F = np.array([_flux(w, k) for k in range(3)]).T     # 5 x 3
                grads = [glam[i] for i in range(4)] + [lam[a_] * glam[b_] + lam[b_] * glam[a_] for a_, b_ in VS]
                for i, gphi in enumerate(grads):
                    ref[:, i] -= jw * (F @ gphi)
    err = np.abs(lim - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


def test_abs_flux_jacobians_are_the_absolute_preconditioned_jacobians(OT):
    """|A|, |B| (setup_preconditioner.cpp:289-354) are the absolute values of the
    Weiss-Smith-preconditioned primitive flux Jacobians P dpdc dF/dw: same square,
    commuting, non-negative eigenvalues.  |C| too when |v| = |w|; otherwise its
    eigenvector matrix carries temp2 in entry (4, 1) (:391), kept as the reference has
    it, and |C|^2 differs from (P dpdc dG/dw)^2."""
    m = TM.unit_cube(1)
    O = _mk(OT, m, 0)
    rng = np.random.default_rng(0)
    for trial in range(4):
        w = np.array([1.0 + rng.random(), *(0.3 * rng.standard_normal(3)), 1.0 + rng.random()])
        w[3] = w[2] if trial % 2 == 0 else -w[2]            # |v| = |w|
        b2 = 0.3 + 0.7 * rng.random()
        rho, gogm1 = w[0] / w[4], GAM / (GAM - 1)
        Pws = np.eye(5)
        Pws[0, 0], Pws[4, 0] = b2, (b2 - 1) / (gogm1 * rho)
        dUdw = _jac(_cons, w)
        X3 = O.abs_jacobians(w, b2)
        for k in range(3):
            M = Pws @ np.linalg.solve(dUdw, _jac(lambda x: _flux(x, k), w))
            X = X3[k]
            sc = np.abs(M @ M).max()
            assert np.abs(X @ X - M @ M).max() < 1e-8 * sc
            assert np.abs(X @ M - M @ X).max() < 1e-8 * sc
            assert np.allclose(np.sort(np.linalg.eigvals(X).real), np.sort(np.abs(np.linalg.eigvals(M).real)), atol=1e-8)
    w = np.array([1.2, 0.1, 0.3, -0.05, 1.1])
    X = O.abs_jacobians(w, 0.5)[2]
    rho = w[0] / w[4]
    Pws = np.eye(5)
    Pws[0, 0], Pws[4, 0] = 0.5, (0.5 - 1) / (GAM / (GAM - 1) * rho)
    M = Pws @ np.linalg.solve(_jac(_cons, w), _jac(lambda x: _flux(x, 2), w))
    assert np.abs(X @ X - M @ M).max() > 1e-6 * np.abs(M @ M).max()


def test_tau_and_timestep_without_preconditioning(OT):
    """At Mach >= 0.8 the Weiss-Smith factor is 1 (P = I); inviscid, Tinv = 2 (|A| + |B|
    + |C|) / h, so Tinv Tau = P dpdc = (dU/dw)^-1, and the time-step factor is
    3 (|u| + c) / h."""
    m = TM.unit_cube(1)
    O = _mk(OT, m, 0, mu=0.0)
    w = np.array([1.3, 1.2, 0.4, 0.4, 1.1])                # |v| = |w|, M ~ 1.05
    h, hmax = 0.37, 0.55
    Pinv, Tau, ts = O.tau_timestep(w, h, hmax)
    assert np.allclose(Pinv, np.eye(5), atol=0, rtol=0)
    A, B, C = O.abs_jacobians(w, 1.0)
    Tinv = 2.0 * (A + B + C) / h
    dpdc = np.linalg.inv(_jac(_cons, w))
    assert np.abs(Tinv @ Tau - dpdc).max() < 1e-8 * np.abs(dpdc).max()
    c = np.sqrt(GAM * w[4])
    assert abs(ts - 3.0 * (np.linalg.norm(w[1:4]) + c) / h) < 1e-13 * ts


def test_switch_variables_maps_conservative_to_primitive(OT):
    """switch_variables(w, dU/dw d) = d for the conservative map U(w) = (rho, rho u,
    rho v, rho w, rho E), E = RT/(gamma-1) + |u|^2/2."""
    m = TM.unit_cube(1)
    O = _mk(OT, m, 0)
    rng = np.random.default_rng(3)
    for _ in range(3):
        w = np.array([1.0 + rng.random(), *rng.standard_normal(3), 1.0 + rng.random()])
        d = rng.standard_normal(5)
        out = O.switch_variables(w, _jac(_cons, w) @ d)
        assert np.abs(out - d).max() < 1e-8 * np.abs(d).max()


def test_relaxation_converges(OT):
    """The multistage relaxation with the Weiss-Smith preconditioner drives the residual
    of a perturbed subsonic state to round-off (p = 1, 3^3 cubes)."""
    m = TM.unit_cube(3, 0.2, 1)
    O = _mk(OT, m, 0, mu=5e-2)
    v, e = _random_state(m, 0, 3, amp=0.02)
    bv = m.boundary_vertex_mask()
    v[bv, 1:] = STATE[1:]
    O.set_state(v, e)
    for it in range(400):
        assert O.setup_preconditioner() == 0
        O.update()
    O.rsdl()
    rv, _ = O.get_res()
    rv[bv, 1:] = 0
    assert np.abs(rv).max() < 1e-12


def test_viscous_weak_form_against_an_independent_quadrature(OT):
    """With tau = 0, stage 0 (beta = 1): lf_re(n, i) = +int_T grad(phi_i) . Fv_n dV with the
    Newtonian stress tau_ij = mu (du_i/dx_j + du_j/dx_i - 2/3 delta_ij div u) for the
    momentum rows and tau.u + (kcond / R) grad(RT) for energy, kcond = R mu gamma /
    (prandtl (gamma - 1)) (tet_hp/cns/allocate.cpp:50); linear u and RT make the stress
    and heat flux constant, so both rules are exact."""
    m = TM.unit_cube(1, 0.0, 0)
    mu, pr = 0.03, 0.72
    O = OT.TetHpCns(m, log2p=1, mu=mu, gamma=GAM, prandtl=pr, R=1.0, bc=(0.0, *STATE[1:]))
    rng = np.random.default_rng(11)
    t = 4
    X = m.pnts[m.tet_pnt[t]]
    wv = np.tile(STATE, (4, 1)) + 0.2 * rng.standard_normal((4, 5))
    wv[:, 0] = 1.0
    wv[:, 4] = 1.0 + 0.2 * rng.random(4)
    uht = np.zeros((5, 10))
    uht[:, :4] = wv.T
    O.set_tau(np.zeros((m.ntet, 5, 5)))
    lre, _ = O.element_rsdl(t, 0, uht)
    Ti = np.linalg.inv(np.vstack([X.T, np.ones(4)]))
    glam = Ti[:, :3]
    vol = abs(np.linalg.det(np.vstack([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6.0)
    gu = np.array([glam.T @ wv[:, 1 + i] for i in range(3)])       # du_i / dx_j
    grt = glam.T @ wv[:, 4]
    stress = mu * (gu + gu.T - 2.0 / 3.0 * np.trace(gu) * np.eye(3))
    kc = 1.0 * mu / pr * GAM / (GAM - 1.0)
    xg, wg = np.polynomial.legendre.leggauss(6)
    xg, wg = 0.5 * (xg + 1), 0.5 * wg
    ref = np.zeros((5, 10))
    for a, wa in zip(xg, wg):
        for b, wb in zip(xg, wg):
            for c, wc in zip(xg, wg):
                l1, l2, l3 = a, b * (1 - a), c * (1 - a) * (1 - b)
                lam = np.array([1 - l1 - l2 - l3, l1, l2, l3])
                jw = wa * wb * wc * (1 - a) ** 2 * (1 - b) * 6.0 * vol
                u = lam @ wv[:, 1:4]
                Fv = np.zeros((5, 3))
                Fv[1:4] = stress
                Fv[4] = stress @ u + kc * grt
                grads = [glam[i] for i in range(4)] + [lam[a_] * glam[b_] + lam[b_] * glam[a_] for a_, b_ in VS]
                for i, gphi in enumerate(grads):
                    ref[:, i] += jw * (Fv @ gphi)
    err = np.abs(lre - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err


def test_stabilisation_flux_jacobians(OT):
    """The matrices of the stabilisation term (rsdl.cpp:799-826) are the flux Jacobians
    with respect to the primitive unknowns, dF_k/dw (finite differences)."""
    m = TM.unit_cube(1)
    O = _mk(OT, m, 0)
    rng = np.random.default_rng(5)
    for _ in range(3):
        w = np.array([1.0 + rng.random(), *rng.standard_normal(3), 1.0 + rng.random()])
        for k, X in enumerate(O.flux_jacobians(w)):
            J = _jac(lambda x: _flux(x, k), w)
            assert np.abs(X - J).max() < 1e-8 * np.abs(J).max()
