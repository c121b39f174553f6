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
