"""Pins of the tri_hp_cns oracle (oracle/otri_hp_cns.cpp) by invariants and by
independent constructions (numpy eigen-decompositions, finite differences)."""
import numpy as np
import pytest

INFLOW = 0b1110


def _mesh(n, jitter=0.2, seed=1, essential=INFLOW):
    from hpmp import mesh as M
    return M.unit_square(n, jitter, seed, essential_mask=essential)


def _cns(oracle_mod, m, log2p, bc, mu=1e-2, **kw):
    return oracle_mod.TriHpCns(m, log2p=log2p, mu=mu, bc=bc, **kw)


@pytest.mark.parametrize("log2p", [0, 1, 2])
@pytest.mark.parametrize("stage", [5, 0, 1])
def test_freestream_residual_vanishes(oracle_mod, log2p, stage):
    from hpmp import workloads as W
    m = _mesh(5, 0.2, 3)
    st, bc = W.cns_freestream(m, 1 << log2p, mach=0.5)
    T = _cns(oracle_mod, m, log2p, bc)
    T.set_state(*st)
    for s in range(stage + 1) if stage < 5 else [5]:
        T.rsdl(s)
    for a in T.get_res():
        if a.size:
            assert np.abs(a).max() < 1e-13


def _boundary_flux_integral(m, st, bc, p, nq=12):
    """mass and momentum boundary integrals of the edge flux, independently: on the
    boundary RT = bc RT, so rho_bc u.n = p_bc/RT_bc (u.n) is a polynomial trace."""
    xg, wg = np.polynomial.legendre.leggauss(nq)
    v = st[0]
    tot = np.zeros(3)
    for e in range(m.nebd):
        for j in range(m.ebdry_start[e], m.ebdry_start[e + 1]):
            sind = m.ebdry_seg[j]
            a, b = m.seg_pnt[sind]
            P0, P1 = m.pnts[a], m.pnts[b]
            t = P1 - P0
            L = np.hypot(*t)
            nrm = np.array([t[1], -t[0]]) / L
            for xq, wq in zip(xg, wg):
                l0, l1 = 0.5 * (1 - xq), 0.5 * (1 + xq)
                # u, v, RT have no side modes on the boundary: their traces are linear
                u = v[a] * l0 + v[b] * l1
                un = u[1] * nrm[0] + u[2] * nrm[1]
                f0 = bc[0] / u[3] * un
                tot += wq * 0.5 * L * np.array([f0, f0 * u[1] + bc[0] * nrm[0], f0 * u[2] + bc[0] * nrm[1]])
    return tot


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_mass_and_momentum_conservation(oracle_mod, log2p):
    """Sum over vertices of the continuity and momentum residuals equals the boundary
    integral of the edge flux (interior terms integrate the gradient of a constant)."""
    from hpmp import workloads as W
    p = 1 << log2p
    m = _mesh(5, 0.2, 4)
    st, bc = W.cns_gaussian_state(m, p, seed=3, amp=0.05, width=0.2, noise=0.0)
    T = _cns(oracle_mod, m, log2p, bc)
    T.set_state(*st)
    T.rsdl()
    rv = T.get_res()[0]
    ref = _boundary_flux_integral(m, st, bc, p)
    assert np.allclose(rv[:, :3].sum(axis=0), ref, rtol=0, atol=1e-13 * max(1.0, np.abs(ref).max()))


@pytest.mark.parametrize("log2p", [1, 2])
def test_rotation_covariance(oracle_mod, log2p):
    """Rotating the mesh and the velocity by 90 degrees rotates the momentum residual
    (viscous, heat and stabilisation terms included); mass and energy are unchanged."""
    from hpmp import workloads as W
    from hpmp.mesh import TriMesh
    p = 1 << log2p
    m = _mesh(4, 0.2, 5)
    st, bc = W.cns_gaussian_state(m, p, seed=6, amp=0.05, width=0.2, noise=1e-2)
    R = np.array([[0.0, -1.0], [1.0, 0.0]])
    m2 = TriMesh(**{k: getattr(m, k) for k in m.__dataclass_fields__})
    m2.pnts = m.pnts @ R.T
    bc2 = (bc[0], 0.0, bc[1], bc[3])
    st2 = [a.copy() for a in st]
    for a in st2:
        if a.size:
            u = a[..., 1].copy(); vv = a[..., 2].copy()
            a[..., 1] = -vv; a[..., 2] = u
    T1 = _cns(oracle_mod, m, log2p, bc); T1.set_state(*st); T1.rsdl()
    T2 = _cns(oracle_mod, m2, log2p, bc2); T2.set_state(*st2); T2.rsdl()
    for a, b in zip(T1.get_res(), T2.get_res()):
        if not a.size:
            continue
        rot = a.copy()
        rot[..., 1] = -a[..., 2]; rot[..., 2] = a[..., 1]
        assert np.abs(rot - b).max() < 1e-12 * max(1.0, np.abs(a).max())


def _abs_matrix(M):
    lam, V = np.linalg.eig(M)
    return np.real(V @ np.diag(np.abs(lam)) @ np.linalg.inv(V))


@pytest.mark.parametrize("pvu,h", [((1.0, 0.3, -0.2, 1.0), 0.05), ((1.2, 1.5, 0.4, 0.8), 0.01),
                                   ((0.9, 0.05, 0.02, 1.1), 0.2)])
def test_calculate_tau_matches_eigen_construction(oracle_mod, pvu, h):
    """Tau = Tinv^-1 P with Tinv = 2(|P A_w| + |P B_w| + h P S)/h, the matrix absolute
    values built from numpy eigen-decompositions of the preconditioned primitive flux
    Jacobians (not from the reference's closed-form eigenvectors)."""
    from hpmp import workloads as W
    m = _mesh(2, 0.0, 1)
    gam, mu, Pr, R = 1.4, 1e-2, 0.713, 1.0
    T = _cns(oracle_mod, m, 1, (1.0, 0.5, 0.0, 1.0), mu=mu, gamma=gam, prandtl=Pr, Rgas=R)
    Pinv, Tau, ts = T.calculate_tau(np.array(pvu), h)
    pr, u, v, rt = pvu
    rho, ke, c = pr / rt, 0.5 * (u * u + v * v), np.sqrt(gam * rt)
    gm1, gogm1 = gam - 1, gam / (gam - 1)
    kcond = mu / Pr * R * gam / gm1
    nu, alpha = mu / rho, kcond / (rho * gogm1 * R)
    M = max(np.hypot(u, v) / c, 1e-5)
    b2 = min(M * M / (1 - M * M) + (4 * nu / (h * c)) ** 2 + (2 * alpha / (h * c)) ** 2, 1.0) if M < 0.8 else 1.0
    Ppre = np.eye(4); Ppre[0, 0] = b2; Ppre[3, 0] = (b2 - 1) / (gogm1 * rho)
    assert np.allclose(Pinv, np.linalg.inv(Ppre), rtol=1e-14, atol=1e-14)
    # primitive -> conservative Jacobian and its inverse
    E = rt / gm1 + ke
    U = lambda w: np.array([w[0] / w[3], w[0] / w[3] * w[1], w[0] / w[3] * w[2], w[0] / w[3] * (w[3] / gm1 + 0.5 * (w[1] ** 2 + w[2] ** 2))])
    w0 = np.array(pvu)
    eps = 1e-6
    dcdp = np.stack([(U(w0 + eps * e) - U(w0 - eps * e)) / (2 * eps) for e in np.eye(4)], 1)
    dpdc = np.linalg.inv(dcdp)
    P = Ppre @ dpdc
    # conservative fluxes as functions of primitives; A_w = dF/dw
    F = lambda w: np.array([w[0] / w[3] * w[1], w[0] / w[3] * w[1] ** 2 + w[0], w[0] / w[3] * w[1] * w[2],
                            w[0] / w[3] * w[1] * (gogm1 * w[3] + 0.5 * (w[1] ** 2 + w[2] ** 2))])
    G = lambda w: np.array([w[0] / w[3] * w[2], w[0] / w[3] * w[1] * w[2], w[0] / w[3] * w[2] ** 2 + w[0],
                            w[0] / w[3] * w[2] * (gogm1 * w[3] + 0.5 * (w[1] ** 2 + w[2] ** 2))])
    Aw = np.stack([(F(w0 + eps * e) - F(w0 - eps * e)) / (2 * eps) for e in np.eye(4)], 1)
    Bw = np.stack([(G(w0 + eps * e) - G(w0 - eps * e)) / (2 * eps) for e in np.eye(4)], 1)
    S = P @ np.diag([0, nu / h ** 2, nu / h ** 2, alpha / h ** 2])
    Tinv = 2 * (_abs_matrix(P @ Aw) + _abs_matrix(P @ Bw) + h * S) / h
    Tau_ref = np.linalg.solve(Tinv, P)
    assert np.abs(Tau - Tau_ref).max() < 1e-7 * np.abs(Tau_ref).max()
    # 20 power iterations: within 1e-6 of the spectral radius for these well-separated cases
    assert abs(1 / ts - np.abs(np.linalg.eigvals(Tinv)).max()) < 1e-3 * np.abs(np.linalg.eigvals(Tinv)).max()


def test_power_iteration_spectral_radius(oracle_mod):
    A = np.diag([5.0, -1.0, 0.5, 0.1])
    Q, _ = np.linalg.qr(np.random.default_rng(0).standard_normal((4, 4)))
    assert abs(oracle_mod.spectral_radius4(Q @ A @ Q.T) - 5.0) < 1e-12
    assert abs(oracle_mod.spectral_radius4(np.diag([-3.0, 0.3, 0.2, 0.1])) - 3.0) < 1e-12


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_element_jacobian_matches_central_differences(oracle_mod, log2p):
    from hpmp import workloads as W
    p = 1 << log2p
    m = _mesh(3, 0.2, 7)
    st, bc = W.cns_gaussian_state(m, p, seed=8, amp=0.05, width=0.2, noise=1e-2)
    T = _cns(oracle_mod, m, log2p, bc)
    T.set_state(*st)
    K = T.element_jacobian(4, 1e-7, 1e-9)
    # central differences of element_rsdl along a random direction
    tm = (p + 1) * (p + 2) // 2
    T.rsdl()
    rng = np.random.default_rng(9)
    uht = np.zeros((4, tm))
    # gather the element's coefficients through element_rsdl's own input convention
    v, s, i = st
    tp, ts_, sg = m.tri_pnt[4], m.tri_seg[4], m.tri_sgn[4]
    for k in range(3):
        uht[:, k] = v[tp[k]]
    cnt = 3
    for k in range(3):
        for mm in range(p - 1):
            uht[:, cnt] = (sg[k] ** mm) * s[ts_[k], mm]
            cnt += 1
    for k in range(i.shape[1]):
        uht[:, cnt] = i[4, k]
        cnt += 1
    d = rng.standard_normal((4, tm)) * 1e-2
    eps = 1e-5
    lp = sum(T.element_rsdl(4, 0, uht + eps * d))
    lm = sum(T.element_rsdl(4, 0, uht - eps * d))
    fd = ((lp - lm) / (2 * eps)).T.ravel()
    jd = K @ d.T.ravel()
    assert np.abs(fd - jd).max() < 1e-5 * max(1.0, np.abs(fd).max())


def test_switch_variables_is_the_primitive_jacobian(oracle_mod):
    """project_new_variables at p = 1 (vertices only): res <- L^-1 dpdc res, with dpdc
    the derivative of the conservative -> primitive map (finite differences here)."""
    from hpmp import workloads as W
    m = _mesh(3, 0.1, 2)
    st, bc = W.cns_gaussian_state(m, 1, seed=5, amp=0.1, width=0.3)
    T = _cns(oracle_mod, m, 0, bc)
    T.set_state(*st)
    assert T.setup_preconditioner() == 0
    vp, sp, tp, vpc, spc, tpc = T.get_prcn()
    rng = np.random.default_rng(1)
    r = rng.standard_normal((m.npnt, 4))
    T.set_res(r, np.zeros((m.nseg, 0, 4)), np.zeros((m.ntri, 0, 4)))
    T.project_new_variables()
    out = T.get_res()[0]
    gm1 = 0.4

    def prim(Uc):
        rho = Uc[0]; u = Uc[1] / rho; v = Uc[2] / rho
        rt = gm1 * (Uc[3] / rho - 0.5 * (u * u + v * v))
        return np.array([rho * rt, u, v, rt])

    bv = m.boundary_vertex_mask()
    for k in range(m.npnt):
        w = st[0][k]
        rho = w[0] / w[3]
        Uc = np.array([rho, rho * w[1], rho * w[2], rho * (w[3] / gm1 + 0.5 * (w[1] ** 2 + w[2] ** 2))])
        eps = 1e-6
        J = np.stack([(prim(Uc + eps * e) - prim(Uc - eps * e)) / (2 * eps) for e in np.eye(4)], 1)
        ref = np.linalg.solve(np.tril(vpc[k]), J @ r[k])
        if bv[k]:
            ref[1:] = 0.0
        assert np.abs(out[k] - ref).max() < 1e-7 * max(1.0, np.abs(ref).max())


def _dpdc_independent(w, gam):
    """d(primitive)/d(conservative) at primitive state w = (p, u, v, RT), as the
    inverse of the complex-step Jacobian (exact to rounding) of the primitive ->
    conservative map (rho = p/RT, rho u, rho v, rho (RT/(gam-1) + |u|^2/2))."""
    gm1 = gam - 1.0

    def cons(x):
        rho = x[0] / x[3]
        return np.array([rho, rho * x[1], rho * x[2], rho * (x[3] / gm1 + 0.5 * (x[1] ** 2 + x[2] ** 2))])

    h = 1e-30
    w = np.asarray(w, dtype=complex)
    J = np.stack([cons(w + 1j * h * e).imag / h for e in np.eye(4)], 1)
    return np.linalg.inv(J)


@pytest.mark.parametrize("log2p", [1, 2])
def test_project_new_variables_is_exact_for_polynomial_fields(oracle_mod, log2p):
    """project_new_variables (tri_hp/cns/update.cpp:5-245) maps the residual field r(x)
    to L(x)^-1 dpdc(u(x)) r(x): exactly at the vertices, by 1-D L2 projection on the
    sides (vertex values held) and by 2-D L2 projection on the interior modes.  With a
    uniform state and the same lower-triangular block L at every vertex the target is
    C r(x), C = L^-1 dpdc, a polynomial of the element's own degree, so both
    projections must reproduce it: every coefficient (vertex, side mode, interior mode)
    becomes C times its old value.  A wrong sign of the odd side modes, a wrong mode
    index, or a side / interior mass matrix that is not the exact one all fail this."""
    p = 1 << log2p
    m = _mesh(4, 0.25, 11, essential=0)   # no essential boundaries: no Dirichlet zeroing
    w = np.array([1.3, 0.4, -0.25, 0.9])
    T = _cns(oracle_mod, m, log2p, (1.0, 0.5, 0.0, 1.0))
    st = [np.zeros(sh) for sh in T.shapes()]
    st[0][:] = w
    T.set_state(*st)
    rng = np.random.default_rng(17 + log2p)
    Lfull = rng.standard_normal((4, 4)) * 0.3 + np.diag([2.0, 1.5, 1.8, 2.2])
    T.set_vpc(np.broadcast_to(Lfull, (m.npnt, 4, 4)).copy())   # only the lower triangle may be used
    r = [rng.standard_normal(sh) for sh in T.shapes()]
    T.set_res(*r)
    T.project_new_variables()
    out = T.get_res()
    C = np.linalg.solve(np.tril(Lfull), _dpdc_independent(w, 1.4))
    for a, b in zip(out, r):
        if a.size:
            ref = b @ C.T
            assert np.abs(a - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_inflow_modify_boundary_residual_closed_form(oracle_mod, log2p):
    """inflow::modify_boundary_residual (tri_hp/cns/bdry.cpp:247-340): with u, v, RT
    prescribed, the conservative momentum and energy residuals follow from the mass
    residual as rho u = u rho, rho v = v rho, rho E = (RT/(gam-1) + |u|^2/2) rho.  On
    the inflow vertices that is pointwise; on the inflow sides the side modes are the
    1-D L2 projection of b_n r_mass(x) with the vertex values held, which for a
    constant boundary state is exactly b_n times the mass residual's side modes.
    Nothing off the inflow boundary changes."""
    p = 1 << log2p
    gam = 1.4
    m = _mesh(4, 0.25, 12)               # every boundary edge is an inflow edge
    bc = (1.1, 0.6, -0.2, 0.8)
    T = _cns(oracle_mod, m, log2p, bc, gamma=gam)
    rng = np.random.default_rng(23 + log2p)
    r = [rng.standard_normal(sh) for sh in T.shapes()]
    T.set_res(*r)
    T.modify_boundary_residual()
    ov, os_, oi = T.get_res()
    b = np.array([1.0, bc[1], bc[2], bc[3] / (gam - 1) + 0.5 * (bc[1] ** 2 + bc[2] ** 2)])
    bv, bs = m.boundary_vertex_mask(), m.boundary_segment_mask()
    ev = r[0].copy()
    ev[bv, 1:] = r[0][bv, :1] * b[1:]
    assert np.abs(ov - ev).max() < 1e-14 * np.abs(ev).max()
    if p > 1:
        es = r[1].copy()
        es[bs, :, 1:] = r[1][bs, :, :1] * b[1:]
        assert np.abs(os_ - es).max() < 1e-12 * np.abs(es).max()
    if oi.size:
        assert np.array_equal(oi, r[2])


@pytest.mark.parametrize("mu", [2e-2, 1e-3])
@pytest.mark.parametrize("log2p", [1, 2])
def test_residual_mesh_scaling_law(oracle_mod, log2p, mu):
    """Scaling the mesh by s at fixed Reynolds and Peclet numbers (mu -> s mu, so kcond
    -> s kcond) and unchanged state multiplies every residual coefficient by s
    (fluxes times d(area)/d(length) = s), stabilisation included: the per-point tau
    (Tinv = 2(|A| + |B| + h S)/h with S ~ mu/h^2) must scale as a length.  A wrong power
    of h in Tinv, in S, in the low-Mach parameter b2 (live at mu = 1e-3, saturated at 1
    at mu = 2e-2) or in the element size fails this; the constants do not."""
    from hpmp import workloads as W
    from hpmp.mesh import TriMesh
    p = 1 << log2p
    m = _mesh(4, 0.2, 13)
    st, bc = W.cns_gaussian_state(m, p, seed=4, amp=0.05, width=0.2, noise=1e-2)
    out = []
    for s in (1.0, 3.0):
        ms = TriMesh(**{**m.__dict__, "pnts": m.pnts * s})
        T = _cns(oracle_mod, ms, log2p, bc, mu=mu * s)
        T.set_state(*st)
        T.rsdl()
        out.append(T.get_res())
    for a, b in zip(out[1], out[0]):
        if a.size:
            assert np.abs(a - 3.0 * b).max() < 1e-12 * np.abs(3.0 * b).max()


@pytest.mark.parametrize("log2p", [1, 2])
def test_project_new_variables_brute_force(oracle_mod, log2p):
    """project_new_variables (tri_hp/cns/update.cpp:5-245) on a non-uniform state with
    a different block L at every vertex, against an independent restatement: the
    target field L(x)^-1 dpdc(u(x)) r(x) (L interpolated linearly, only its lower
    triangle used, dpdc by complex step) evaluated with the closed-form basis at the
    basis' own Gauss points (Gauss-Legendre on sides and in xi, Gauss-Jacobi(1,0) in
    eta, from scipy), then the discrete L2 projection onto the side modes (vertex
    values held) and onto the interior modes (vertex and side values held).  It pins
    the sign of the odd side modes in the element gather, the mode ordering, and
    which vertex state and block the side and interior targets see."""
    from scipy.special import eval_jacobi, roots_jacobi
    from hpmp import workloads as W
    from test_oracle_basis import _basis_closed_form
    p = 1 << log2p
    sm, im = p - 1, (p - 1) * (p - 2) // 2
    gam = 1.4
    m = _mesh(3, 0.25, 31, essential=0)
    st, bc = W.cns_gaussian_state(m, p, seed=6, amp=0.1, width=0.3, noise=3e-2)
    T = _cns(oracle_mod, m, log2p, bc, gamma=gam)
    T.set_state(*st)
    rng = np.random.default_rng(41 + log2p)
    vpc = rng.standard_normal((m.npnt, 4, 4)) * 0.2 + np.diag([2.0, 1.5, 1.8, 2.2])
    T.set_vpc(vpc)
    r = [rng.standard_normal(sh) for sh in T.shapes()]
    T.set_res(*r)
    T.project_new_variables()
    ov, os_, oi = T.get_res()

    def dpdc(w):
        return _dpdc_independent(w, gam)

    def target(Lq, wq, rq):
        return np.linalg.solve(np.tril(Lq), dpdc(wq) @ rq)

    # vertices: pointwise
    for k in range(m.npnt):
        assert np.allclose(ov[k], target(vpc[k], st[0][k], r[0][k]), rtol=0, atol=1e-12 * np.abs(ov).max())

    # sides: 1-D projection with the vertex values held
    xs, ws = np.polynomial.legendre.leggauss(p + 1)

    def side1d(k, t):
        return 0.25 * (1 - t) * (1 + t) * eval_jacobi(k, 2, 2, t)

    nrm = [1.0 / np.sqrt((ws * side1d(k, xs) ** 2).sum()) for k in range(sm)]
    psi = np.array([side1d(k, xs) * nrm[k] for k in range(sm)])          # sm x q
    l0, l1 = 0.5 * (1 - xs), 0.5 * (1 + xs)
    Mss = (psi * ws) @ psi.T
    for sd in range(m.nseg):
        a, b = m.seg_pnt[sd]
        u = np.outer(l0, st[0][a]) + np.outer(l1, st[0][b]) + psi.T @ st[1][sd]
        rr = np.outer(l0, r[0][a]) + np.outer(l1, r[0][b]) + psi.T @ r[1][sd]
        Lq = l0[:, None, None] * vpc[a] + l1[:, None, None] * vpc[b]
        tq = np.array([target(Lq[q], u[q], rr[q]) for q in range(p + 1)])
        held = np.outer(l0, ov[a]) + np.outer(l1, ov[b])
        c = np.linalg.solve(Mss, (psi * ws) @ (tq - held))
        assert np.abs(os_[sd] - c).max() < 1e-12 * np.abs(os_).max()

    if im == 0:
        return
    # interior: 2-D projection with vertex and side values held
    xg, wx = np.polynomial.legendre.leggauss(p + 1)
    eg, we = roots_jacobi(p + 1, 1.0, 0.0)
    X, E = np.meshgrid(xg, eg, indexing="ij")
    Wq = np.outer(wx, 0.5 * we).ravel()
    phi = _basis_closed_form(p, X.ravel(), E.ravel())                     # tm x q
    nb = 3 + 3 * sm
    MII = (phi[nb:] * Wq) @ phi[nb:].T
    for t in range(m.ntri):
        vs, sg, ss = m.tri_pnt[t], m.tri_sgn[t], m.tri_seg[t]

        def gather(v, s, i):
            rows = [v[vs[k]] for k in range(3)]
            for k in range(3):
                rows += [float(sg[k]) ** mm * s[ss[k], mm] for mm in range(sm)]
            return np.array(rows + [i[t, k] for k in range(im)])

        u = phi.T @ gather(*st)
        rr = phi.T @ gather(*r)
        Lq = np.einsum("kq,kab->qab", phi[:3], vpc[vs])
        tq = np.array([target(Lq[q], u[q], rr[q]) for q in range(phi.shape[1])])
        held = phi[:nb].T @ gather(ov, os_, np.zeros_like(oi))[:nb]
        c = np.linalg.solve(MII, (phi[nb:] * Wq) @ (tq - held))
        assert np.abs(oi[t] - c).max() < 1e-11 * np.abs(oi).max()
