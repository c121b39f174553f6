"""Boundary flux types of tri_hp_ins in the oracle: generic (inflow, outflow, symmetry,
tri_hp/ins/bdry_ins.h:34-49), euler (bdry_ins.h:96-113) and characteristic
(tri_hp/ins/bdry.cpp:245-320).

Pins that the mathematics fixes:
* at the freestream (interior state equal to the ibc state) the characteristic swap
  returns that state, so its flux is the euler and the generic flux of the freestream;
* the swap is symmetric: exchanging the interior and freestream states and reversing
  the normal returns the same boundary state, so the flux changes sign (states of equal
  speed, which keep the swap's gam);
* a uniform flow equal to the freestream is steady with every boundary type: its full
  residual vanishes at every mode (the element integrals of a constant flux cancel
  against the boundary fluxes of the same state).
"""
import numpy as np
import pytest

from conftest import rel_err


def _states(rng, n=50):
    for _ in range(n):
        th = rng.uniform(0, 2 * np.pi)
        nrm = np.array([np.cos(th), np.sin(th)])
        u = np.append(rng.standard_normal(2), rng.standard_normal())
        yield u, nrm, rng.uniform(0.01, 0.2), int(rng.choice([1, 2, 4])), rng.uniform(0.5, 2.0), rng.uniform(0.0, 0.1)


def test_characteristic_flux_at_the_freestream_is_its_euler_flux(oracle_mod):
    import oracle
    rng = np.random.default_rng(3)
    for u, nrm, sl, p, rho, mu in _states(rng):
        fc = oracle.edge_flux(2, u, nrm, sl, p, rho, mu, u)
        fe = oracle.edge_flux(1, u, nrm, sl, p, rho, mu, u)
        fg = oracle.edge_flux(0, u, nrm, sl, p, rho, mu, u)
        un = u[:2] @ nrm
        ref = np.array([rho * un * u[0] + u[2] * nrm[0], rho * un * u[1] + u[2] * nrm[1], rho * un])
        assert np.abs(fc - ref).max() < 1e-12 * (1 + np.abs(ref).max())
        assert np.abs(fe - ref).max() < 1e-14 * (1 + np.abs(ref).max())
        assert np.abs(fg - ref).max() < 1e-14 * (1 + np.abs(ref).max())


def test_characteristic_swap_is_symmetric(oracle_mod):
    import oracle
    rng = np.random.default_rng(4)
    for u, nrm, sl, p, rho, mu in _states(rng):
        th = rng.uniform(0, 2 * np.pi)
        sp = np.linalg.norm(u[:2])
        ub = np.array([sp * np.cos(th), sp * np.sin(th), rng.standard_normal()])
        f = oracle.edge_flux(2, u, nrm, sl, p, rho, mu, ub)
        g = oracle.edge_flux(2, ub, -nrm, sl, p, rho, mu, u)
        assert np.abs(f + g).max() < 1e-12 * (1 + np.abs(f).max())
        # and it differs from both one-sided fluxes (the swap mixes the two states)
        assert np.abs(f - oracle.edge_flux(1, u, nrm, sl, p, rho, mu, ub)).max() > 1e-6


def typed_square(n, jitter=0.2, seed=1):
    """Unit square with one boundary of each kind: bottom symmetry (v essential),
    right outflow (nothing essential), top characteristic, left inflow (u, v)."""
    from hpmp import mesh as M
    m = M.unit_square(n, jitter, seed)
    m.ebdry_essential[:] = [0b010, 0b000, 0b000, 0b011]
    types = [0, 0, 2, 0]
    return m, types


@pytest.mark.parametrize("log2p", [0, 1, 2])
@pytest.mark.parametrize("types", [[0, 0, 2, 0], [0, 1, 2, 1], [2, 2, 2, 2]])
def test_uniform_freestream_is_steady_with_every_boundary_type(oracle_mod, log2p, types):
    m, _ = typed_square(5)
    p = 1 << log2p
    sm, im = p - 1, (p - 1) * (p - 2) // 2
    ub = np.array([0.8, 0.3, 0.25])
    v = np.tile(ub, (m.npnt, 1))
    s = np.zeros((m.nseg, sm, 3))
    i = np.zeros((m.ntri, im, 3))
    O = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    O.set_edge_flux(types, ub)
    O.set_state(v, s, i)
    O.rsdl()
    r = max(np.abs(x).max() for x in O.get_res() if x.size)
    assert r < 1e-14
    # a different freestream in the ibc leaves a residual on the characteristic and
    # euler boundaries (and the generic one through its pressure)
    O.set_edge_flux(types, ub + [0.1, -0.05, 0.2])
    O.rsdl()
    assert max(np.abs(x).max() for x in O.get_res() if x.size) > 1e-3
