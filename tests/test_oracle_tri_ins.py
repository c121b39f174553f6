"""Pins of the oracle's tri_hp_ins restatement: physical-space weak forms by
brute-force quadrature, discrete conservation, rotation covariance, the lumped
mass inverse on linear data, and the multigrid source identity."""
import numpy as np
import pytest

from test_oracle_basis import _basis_closed_form, _tri_rule


def _mesh(n, jitter=0.2, seed=1, essential=0b011):
    from hpmp import mesh as M
    return M.unit_square(n, jitter, seed, essential_mask=essential)


def _phys_basis(p, verts):
    """phi, grad phi at a high-order rule on the physical triangle (FD in r, s)."""
    X, E, R, S, Wq = _tri_rule(10)
    h = 1e-6

    def phi_rs(r, s):
        return _basis_closed_form(p, 2 * (1 + r) / (1 - s) - 1, s)

    ph = phi_rs(R, S)
    dpr = (phi_rs(R + h, S) - phi_rs(R - h, S)) / (2 * h)
    dps = (phi_rs(R, S + h) - phi_rs(R, S - h)) / (2 * h)
    v = verts
    J = 0.5 * np.array([[v[2, 0] - v[1, 0], v[0, 0] - v[1, 0]], [v[2, 1] - v[1, 1], v[0, 1] - v[1, 1]]])
    det = np.linalg.det(J)
    Jit = np.linalg.inv(J).T
    gx = Jit[0, 0] * dpr + Jit[0, 1] * dps
    gy = Jit[1, 0] * dpr + Jit[1, 1] * dps
    return ph, gx, gy, Wq * det


def _low_order_coeffs(p, rng, maxorder):
    from hpmp.workloads import mode_orders
    side_k, int_k = mode_orders(p)
    order = np.concatenate([[1, 1, 1], np.tile(side_k, 3), int_k])
    u = rng.standard_normal((3, len(order))) * 0.3
    u[0, :3] += 1.0
    u[:, order > maxorder] = 0.0
    return u


@pytest.mark.parametrize("log2p,maxorder", [(0, 1), (1, 2), (2, 3)])
def test_inviscid_residual_is_the_weak_form(oracle_mod, log2p, maxorder):
    """lf_im = -int grad(phi) . F(u) dA, F the Euler flux + pressure, for states whose
    integrands the reference's Gauss rule integrates exactly."""
    p = 1 << log2p
    m = _mesh(3, 0.25, 7)
    T = oracle_mod.TriHpIns(m, log2p=log2p, mu=0.0)
    rng = np.random.default_rng(log2p)
    for t in (0, 5, 11):
        uht = _low_order_coeffs(p, rng, maxorder)
        ph, gx, gy, w = _phys_basis(p, m.pnts[m.tri_pnt[t]])
        u, v, pr = uht @ ph
        flux = {0: (u * u + pr, u * v), 1: (v * u, v * v + pr), 2: (u, v)}
        ref = np.array([-((gx * flux[n][0] + gy * flux[n][1]) * w).sum(axis=1) for n in range(3)])
        _, li = T.element_rsdl(t, 1, uht)
        assert np.abs(li - ref).max() < 1e-8 * np.abs(ref).max()


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_viscous_residual_is_the_weak_form(oracle_mod, log2p):
    """With the stabilisation off (dissipation 0), lf_re = +int grad(phi) : mu (grad u + grad u^T)."""
    p = 1 << log2p
    mu = 0.03
    m = _mesh(3, 0.25, 7)
    T = oracle_mod.TriHpIns(m, log2p=log2p, mu=mu, adis=0.0)
    rng = np.random.default_rng(10 + log2p)
    for t in (1, 8):
        uht = _low_order_coeffs(p, rng, 99)
        ph, gx, gy, w = _phys_basis(p, m.pnts[m.tri_pnt[t]])
        ux, uy = uht @ gx, uht @ gy
        tau = [[2 * ux[0], uy[0] + ux[1]], [ux[1] + uy[0], 2 * uy[1]]]
        ref = np.array([((gx * mu * tau[n][0] + gy * mu * tau[n][1]) * w).sum(axis=1) for n in range(2)])
        lr, _ = T.element_rsdl(t, 0, uht)
        assert np.abs(lr[:2] - ref).max() < 1e-8 * np.abs(ref).max()
        assert np.abs(lr[2]).max() == 0.0


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_freestream_residual_vanishes(oracle_mod, log2p):
    from hpmp import workloads as W
    m = _mesh(6, 0.2, 3)
    T = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    T.set_state(*W.freestream(m, 1 << log2p, 0.8, -0.3))
    T.rsdl()
    for a in T.get_res():
        if a.size:
            assert np.abs(a).max() < 1e-14


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_constant_pressure_is_a_null_mode(oracle_mod, log2p):
    """With u, v prescribed on every boundary edge, adding a constant to the pressure
    changes no residual row that survives the Dirichlet conditions."""
    from hpmp import workloads as W
    m = _mesh(6, 0.2, 5)
    T = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    st = W.noisy_freestream(m, 1 << log2p, seed=3)
    T.set_state(*st)
    T.rsdl()
    r1 = [a.copy() for a in T.get_res()]
    v = st[0].copy()
    v[:, 2] += 0.37
    T.set_state(v, st[1], st[2])
    T.rsdl()
    r2 = T.get_res()
    bv, bs = m.boundary_vertex_mask(), m.boundary_segment_mask()
    for k, (a, b) in enumerate(zip(r1, r2)):
        if not a.size:
            continue
        d = np.abs(a - b)
        if k == 0:
            d[bv, :2] = 0
        elif k == 1:
            d[bs, :, :2] = 0
        assert d.max() < 1e-13 * max(1.0, np.abs(a).max())


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_continuity_conservation(oracle_mod, log2p):
    """sum over vertices of the continuity residual = boundary integral of rho u.n,
    computed independently along every boundary edge with a 12-point rule."""
    from hpmp import workloads as W
    p = 1 << log2p
    m = _mesh(5, 0.2, 4, essential=0)
    T = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    v, s, i = W.noisy_freestream(m, p, seed=2)
    rng = np.random.default_rng(9)
    bv = m.boundary_vertex_mask()
    v[bv, :2] += 0.1 * rng.standard_normal((bv.sum(), 2))
    if s.size:
        bs = m.boundary_segment_mask()
        s[bs, :, :2] += 0.05 * rng.standard_normal((bs.sum(), p - 1, 2))
    T.set_state(v, s, i)
    T.rsdl()
    rv, _, _ = T.get_res()
    xg, wg = np.polynomial.legendre.leggauss(12)
    flux = 0.0
    from scipy.special import eval_jacobi
    xs, ws = np.polynomial.legendre.leggauss(p + 1)
    for sd in m.ebdry_seg:
        a, b = m.seg_pnt[sd]
        pa, pb = m.pnts[a], m.pnts[b]
        tvec = pb - pa
        nrm = np.array([tvec[1], -tvec[0]])  # outward * length (domain on the left)
        ut = np.outer(0.5 * (1 - xg), v[a, :2]) + np.outer(0.5 * (1 + xg), v[b, :2])
        for k in range(p - 1):
            f = 0.25 * (1 - xg) * (1 + xg) * eval_jacobi(k, 2, 2, xg)
            fn = 1.0 / np.sqrt((ws * (0.25 * (1 - xs) * (1 + xs) * eval_jacobi(k, 2, 2, xs)) ** 2).sum())
            ut += np.outer(f * fn, s[sd, k, :2])
        flux += 0.5 * (wg * (ut @ nrm)).sum()
    assert abs(rv[:, 2].sum() - flux) < 1e-12 * max(1.0, abs(flux))


@pytest.mark.parametrize("log2p", [1, 2])
def test_rotation_covariance(oracle_mod, log2p):
    """Rotating the mesh and the velocity rotates the momentum residual and leaves the
    continuity residual unchanged (full residual: stabilisation and edge fluxes)."""
    from hpmp import workloads as W
    p = 1 << log2p
    m = _mesh(4, 0.2, 8)
    v, s, i = W.noisy_freestream(m, p, seed=4)
    T = oracle_mod.TriHpIns(m, log2p=log2p, mu=2e-2)
    T.set_state(v, s, i)
    T.rsdl()
    r0 = T.get_res()
    th = 0.7
    Rm = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    m2 = _mesh(4, 0.2, 8)
    m2.pnts = m.pnts @ Rm.T
    T2 = oracle_mod.TriHpIns(m2, log2p=log2p, mu=2e-2)
    rot = lambda a: np.concatenate([a[..., :2] @ Rm.T, a[..., 2:]], axis=-1)
    T2.set_state(rot(v), rot(s), rot(i))
    T2.rsdl()
    r1 = T2.get_res()
    scale = max(np.abs(a).max() for a in r0 if a.size)
    for a, b in zip(r0, r1):
        if a.size:
            assert np.abs(rot(a) - b).max() < 1e-12 * scale


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_minvrt_inverts_mass_on_linear_fields(oracle_mod, log2p):
    """With tprcn = 1 the lumped, statically condensed inverse reproduces a linear
    field (p >= 2) or a constant (p = 1, row-sum lumping) from its mass residual."""
    from hpmp.workloads import mode_orders  # noqa: F401
    p = 1 << log2p
    m = _mesh(6, 0.2, 5, essential=0)
    T = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    B = oracle_mod.Basis(p)
    tm, sm, im = B.tm, B.sm, B.im
    Mref = np.array([B.intgrt(B.proj_val(np.eye(tm)[k])) for k in range(tm)])
    L = 0.3 + (1.7 * m.pnts[:, 0] - 0.6 * m.pnts[:, 1] if p > 1 else 0.0)
    L = np.broadcast_to(L, (m.npnt,)).copy()
    rv = np.zeros((m.npnt, 3)); rs = np.zeros((m.nseg, sm, 3)); ri = np.zeros((m.ntri, im, 3))
    for t in range(m.ntri):
        x = np.zeros(tm); x[:3] = L[m.tri_pnt[t]]
        r = Mref @ x
        rv[m.tri_pnt[t], :] += r[:3, None]
        for sd in range(3):
            sg = m.tri_sgn[t, sd]
            for k in range(sm):
                rs[m.tri_seg[t, sd], k, :] += sg ** k * r[3 + sd * sm + k]
        for k in range(im):
            ri[t, k, :] += r[3 * p + k]
    T.set_tprcn(np.ones((m.ntri, 3)))
    T.assemble_prcn()
    T.set_res(rv, rs, ri)
    T.minvrt()
    ov, os_, oi = T.get_res()
    assert np.abs(ov - L[:, None]).max() < 1e-12
    for a in (os_, oi):
        if a.size:
            assert np.abs(a).max() < 1e-12


def test_update_single_stage_is_explicit_step(oracle_mod):
    """nstage = 1, alpha = beta = 1: update is ug <- ug - cfl * minvrt(R(ug))."""
    from hpmp import workloads as W
    m = _mesh(5, 0.2, 6)
    kw = dict(log2p=2, mu=1e-2, alpha=(1.0,), beta=(1.0,))
    A = oracle_mod.TriHpIns(m, **kw)
    B = oracle_mod.TriHpIns(m, **kw)
    st = W.noisy_freestream(m, 4, seed=1)
    A.set_state(*st); B.set_state(*st)
    A.setup_preconditioner(); B.setup_preconditioner()
    A.update()
    B.rsdl(0); B.minvrt()
    r = B.get_res()
    for a, u0, rr in zip(A.get_state(), st, r):
        if a.size:
            assert np.abs(a - (u0 - 1.0 * rr)).max() < 1e-14 * max(1.0, np.abs(u0).max())


def test_pmg_source_reproduces_restricted_residual(oracle_mod):
    """After mg_restrict the coarse rsdl adds dres = fadd*res0 - R_c, so at the
    unchanged state the coarse residual equals the restricted fine residual."""
    from hpmp import workloads as W
    m = _mesh(5, 0.2, 6)
    T = oracle_mod.TriHpIns(m, log2p=2, mu=1e-2)
    T.set_state(*W.noisy_freestream(m, 4, seed=3))
    T.rsdl()
    fv, fs, _ = T.get_res()
    T.mg_restrict()
    assert T.log2p == 1
    T.rsdl(0)
    cv, cs, _ = T.get_res()
    assert np.abs(cv - fv).max() < 1e-13 * np.abs(fv).max()
    assert np.abs(cs[:, :1] - fs[:, :1]).max() < 1e-13 * np.abs(fs).max()


def test_element_jacobian_matches_central_differences(oracle_mod):
    """The forward-difference element Jacobian (tri_hp/nstage.cpp:131-165) agrees with a
    central-difference Jacobian of element_rsdl to first order."""
    from hpmp import workloads as W
    m = _mesh(3, 0.2, 2)
    T = oracle_mod.TriHpIns(m, log2p=1, mu=1e-2)
    T.set_state(*W.noisy_freestream(m, 2, seed=5))
    K = T.element_jacobian(4)
    # rebuild the element coefficients from the state
    v, s, _ = T.get_state()
    tp, ts, sg = m.tri_pnt[4], m.tri_seg[4], m.tri_sgn[4]
    uht = np.zeros((3, 6))
    uht[:, :3] = v[tp].T
    for k in range(3):
        uht[:, 3 + k] = s[ts[k], 0, :]
    h = 1e-6
    Kc = np.zeros_like(K)
    for mode in range(6):
        for var in range(3):
            up = uht.copy(); up[var, mode] += h
            dn = uht.copy(); dn[var, mode] -= h
            rp = sum(T.element_rsdl(4, 0, up)); rm = sum(T.element_rsdl(4, 0, dn))
            Kc[:, mode * 3 + var] = ((rp - rm) / (2 * h)).T.ravel()
    assert np.abs(K - Kc).max() < 1e-5 * np.abs(Kc).max()


def _scaled(m, s):
    from hpmp.mesh import TriMesh
    return TriMesh(**{**m.__dict__, "pnts": m.pnts * s})


def _scaled_state(st, c, k=1.0):
    """velocity x c, pressure x k c^2 (coefficient arrays; the basis is linear)."""
    out = []
    for a in st:
        b = a.copy()
        if b.size:
            b[..., :2] *= c
            b[..., 2] *= k * c * c
        out.append(b)
    return out


def _rsdl(oracle_mod, m, st, **kw):
    T = oracle_mod.TriHpIns(m, **kw)
    T.set_state(*st)
    T.rsdl()
    return T.get_res()


@pytest.mark.parametrize("log2p", [1, 2])
def test_residual_mesh_scaling_law(oracle_mod, log2p):
    """x -> s x with the Reynolds number held (mu -> s mu) and the state unchanged
    multiplies every residual coefficient by s: Galerkin fluxes carry d(area)/d(length),
    and the stabilisation (tri_hp/ins/rsdl.cpp:377-395) only keeps that if tau scales as
    length/velocity, i.e. h/(cjcb sqrt(3q + (2 nu/h)^2)) with h a length and cjcb an
    area.  A tau built on h^2, on area instead of cjcb, or on nu/h^2 fails; the
    constants 3, 2 and 0.25(p+1)^2 do not (they stay unpinned)."""
    from hpmp import workloads as W
    m = _mesh(5, 0.2, 21)
    st = W.fourier_state(m, 1 << log2p, seed=5)
    r1 = _rsdl(oracle_mod, m, st, log2p=log2p, mu=3e-2)
    r3 = _rsdl(oracle_mod, _scaled(m, 3.0), st, log2p=log2p, mu=9e-2)
    for a, b in zip(r3, r1):
        if a.size:
            assert np.abs(a - 3.0 * b).max() < 1e-12 * np.abs(3.0 * b).max()


@pytest.mark.parametrize("log2p", [1, 2])
def test_residual_density_scaling_law(oracle_mod, log2p):
    """rho -> k rho, mu -> k mu, p -> k p at fixed velocity multiplies every residual
    row by k, stabilisation included: tau depends on the kinematic viscosity
    nu = mu/rho only.  A tau on mu instead of nu fails."""
    from hpmp import workloads as W
    m = _mesh(5, 0.2, 22)
    st = W.fourier_state(m, 1 << log2p, seed=6)
    r1 = _rsdl(oracle_mod, m, st, log2p=log2p, rho=1.0, mu=2e-2)
    r2 = _rsdl(oracle_mod, m, _scaled_state(st, 1.0, 2.5), log2p=log2p, rho=2.5, mu=5e-2)
    for a, b in zip(r2, r1):
        if a.size:
            assert np.abs(a - 2.5 * b).max() < 1e-12 * np.abs(2.5 * b).max()


@pytest.mark.parametrize("log2p", [1, 2])
def test_residual_velocity_scaling_law(oracle_mod, log2p):
    """u -> c u, p -> c^2 p, mu -> c mu on a fixed mesh.  The continuity residual
    (Galerkin and the tau_0 pressure stabilisation) scales as c.  Momentum splits
    into c^2 (fluxes and the tau_0 = h/(cjcb sqrt(gam)) streamline term) plus c times
    the grad-div term, whose tau = sqrt(q) tau_0 (tri_hp/ins/rsdl.cpp:392) is
    velocity-independent.  So R_mom(c) = c^2 A + c B for three values of c, with B != 0
    when the state has divergence and B = 0 when the discrete divergence vanishes
    (a linear solenoidal field).  tau = q tau_0 would put the grad-div term into c^2
    (B = 0 always); tau independent of q would put it into c^0."""
    from hpmp import workloads as W
    p = 1 << log2p
    m = _mesh(5, 0.2, 23, essential=0)
    kw = dict(log2p=log2p, bc_pressure=0.0)
    st = W.fourier_state(m, p, seed=7)
    R = {c: _rsdl(oracle_mod, m, _scaled_state(st, c), mu=2e-2 * c, **kw) for c in (1.0, 2.0, 3.0)}
    for k in range(3):
        if not R[1.0][k].size:
            continue
        r1, r2, r3 = (R[c][k] for c in (1.0, 2.0, 3.0))
        scale = np.abs(r3).max()
        assert np.abs(r3[..., 2] - 3.0 * r1[..., 2]).max() < 1e-12 * scale
        B = (4.0 * r1[..., :2] - r2[..., :2]) / 2.0
        A = r1[..., :2] - B
        assert np.abs(r3[..., :2] - (9.0 * A + 3.0 * B)).max() < 1e-11 * scale
        if k == 0:
            assert np.abs(B).max() > 1e-6 * scale
    # linear solenoidal velocity: u = (a + b x + c y, d + e x - b y), linear pressure
    from hpmp.workloads import mode_orders  # noqa: F401
    v = np.zeros((m.npnt, 3))
    x, y = m.pnts[:, 0], m.pnts[:, 1]
    v[:, 0] = 0.7 + 0.3 * x - 0.4 * y
    v[:, 1] = -0.2 + 0.5 * x - 0.3 * y
    v[:, 2] = 0.1 * x + 0.2 * y
    lin = [v, np.zeros((m.nseg, p - 1, 3)), np.zeros((m.ntri, (p - 1) * (p - 2) // 2, 3))]
    R = {c: _rsdl(oracle_mod, m, _scaled_state(lin, c), mu=2e-2 * c, **kw) for c in (1.0, 3.0)}
    for a, b in zip(R[3.0], R[1.0]):
        if a.size:
            scale = np.abs(a).max()
            assert np.abs(a[..., :2] - 9.0 * b[..., :2]).max() < 1e-12 * scale


@pytest.mark.parametrize("log2p", [1, 2])
def test_setup_preconditioner_scaling_laws(oracle_mod, log2p):
    """tprcn of tri_hp_ins::setup_preconditioner (tri_hp/ins/setup_preconditioner.cpp:
    166-204) is rho * area * (viscous rate + wave rate) for u, v and that over gam for p:
    under x -> s x, mu -> s mu it scales as s (area s^2 times rates 1/s); under
    u -> c u, mu -> c mu as (c, c, 1/c); under rho -> k rho, mu -> k mu as (k, k, 1)."""
    from hpmp import workloads as W
    p = 1 << log2p
    m = _mesh(5, 0.2, 24)
    st = W.fourier_state(m, p, seed=8)

    def tp(mesh, state, **kw):
        T = oracle_mod.TriHpIns(mesh, log2p=log2p, **kw)
        T.set_state(*state)
        assert T.setup_preconditioner() == 0
        return T.get_prcn()[2]

    t1 = tp(m, st, mu=2e-2)
    ts = tp(_scaled(m, 2.0), st, mu=4e-2)
    assert np.abs(ts - 2.0 * t1).max() < 1e-12 * np.abs(ts).max()
    tc = tp(m, _scaled_state(st, 3.0), mu=6e-2)
    ref = t1 * np.array([3.0, 3.0, 1.0 / 3.0])
    assert np.abs(tc - ref).max() < 1e-12 * np.abs(ref).max()
    tk = tp(m, _scaled_state(st, 1.0, 2.0), rho=2.0, mu=4e-2)
    ref = t1 * np.array([2.0, 2.0, 1.0])
    assert np.abs(tk - ref).max() < 1e-12 * np.abs(ref).max()
