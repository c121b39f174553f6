"""FP32 defect correction of the tri cns, tet ins and tet cns paths.

The FP32 element residual of the inner solve (FP64 state, FP32 arithmetic, FP64 boundary
fluxes and assembly) is pinned against the oracle's FP64 residual at FP32 accuracy; the
defect-corrected relaxation must converge (FP64 judge): tet ins and cns to the oracle's
FP64 solution to 1e-10 (modulo the pressure level), tri cns to a state at which the
oracle's FP64 residual vanishes to 1e-10 (its steady problem has a family of solutions).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu
INFLOW = 0b1110
# FP32 evaluation of the element residual: rounding of the state, of the operators and of
# every product, FP32 accumulation; relative to the largest residual entry
FP32_TOL = 5e-5


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _cns_pair(oracle_mod, n, log2p, mu=1e-2, seed=3, cfl=(2.5, 1.5, 1.0)):
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpCns
    m = M.unit_square(n, 0.2, seed, essential_mask=INFLOW)
    st, bc = W.cns_gaussian_state(m, 1 << log2p, seed=11, amp=0.05, width=0.2, noise=2e-2)
    G = TriHpCns(m, log2p=log2p, mu=mu, bc=bc, cfl=cfl)
    O = oracle_mod.TriHpCns(m, log2p=log2p, mu=mu, bc=bc, cfl=cfl)
    G.set_state(*st)
    O.set_state(*st)
    return m, st, G, O


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_cns_fp32_residual_matches_oracle(oracle_mod, log2p):
    _case(f"cns FP32 residual vs oracle log2p={log2p}")
    m, st, G, O = _cns_pair(oracle_mod, 8, log2p)
    r32 = G.residual_fp32()
    O.rsdl()
    ro = O.get_res()
    err = rel_err(r32, ro)
    print("rel err", err, flush=True)
    assert err < FP32_TOL
    G.rsdl()
    assert rel_err(G.get_residual(), ro) < 1e-12


def _tet_pair(physics, n, log2p, mu, seed):
    import oracle_tet as OT
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpCns, TetHpIns
    m = TM.unit_cube(n, 0.2, seed)
    em0 = (1 << log2p) - 1
    rng = np.random.default_rng(seed + 10)
    bv, be = m.boundary_vertex_mask(), m.boundary_edge_mask()
    if physics == "ins":
        base = np.array([1.0, 0.3, -0.2, 0.5])
        v = base + 0.05 * rng.standard_normal((m.npnt, 4))
        e = 0.05 * rng.standard_normal((m.nseg, em0, 4))
        v[bv, :3] = base[:3]
        e[be, :, :3] = 0.0
        G = TetHpIns(m, log2p=log2p, mu=mu)
        O = OT.TetHpIns(m, log2p=log2p, mu=mu)
        ess = slice(0, 3)
    else:
        base = np.array([1.0, 0.3, 0.1, -0.05, 1.0])
        kw = dict(gamma=1.4, prandtl=0.72, R=1.0)
        v = base + 0.02 * rng.standard_normal((m.npnt, 5))
        e = 0.02 * rng.standard_normal((m.nseg, em0, 5))
        v[bv, 1:] = base[1:]
        e[be, :, 1:] = 0.0
        G = TetHpCns(m, log2p=log2p, mu=mu, bc=tuple(base[1:]), **kw)
        O = OT.TetHpCns(m, log2p=log2p, mu=mu, bc=(0.0, *base[1:]), **kw)
        ess = slice(1, 5)
    G.set_state(v, e)
    O.set_state(v, e)
    return m, (v, e), G, O, ess


@pytest.mark.parametrize("physics", ["ins", "cns"])
@pytest.mark.parametrize("log2p", [0, 1])
def test_tet_fp32_residual_matches_oracle(oracle_mod, physics, log2p):
    _case(f"tet {physics} FP32 residual vs oracle log2p={log2p}")
    m, st, G, O, _ = _tet_pair(physics, 3, log2p, 1e-2, 4)
    G.setup_preconditioner()   # the frozen element tau the residual uses
    O.setup_preconditioner()
    r32 = G.residual_fp32()
    O.rsdl()
    ro = O.get_res()
    err = rel_err(r32, ro)
    print("rel err", err, flush=True)
    assert err < FP32_TOL
    G.rsdl()
    assert rel_err(G.get_residual(), ro) < 1e-12


@pytest.mark.parametrize("physics", ["ins", "cns"])
def test_tet_mixed_reaches_fp64_solution(oracle_mod, physics):
    _case(f"tet {physics} FP32 defect correction converges to the FP64 solution 3^3 p=2")
    from hpmp import _lib
    m, st, G, O, ess = _tet_pair(physics, 3, 1, 5e-2, 2)
    bv, be = m.boundary_vertex_mask(), m.boundary_edge_mask()

    def judge(X, get_res):
        X.rsdl()
        rv, re = get_res()
        rv[bv, ess] = 0
        re[be, :, ess] = 0
        return max(np.abs(rv).max(), np.abs(re).max())

    G.set_precision(_lib.HPG_FP32_DEFECT, 1)
    r0 = judge(G, G.get_residual)
    for it in range(4000):
        assert G.setup_preconditioner() == 0
        G.update()
        if it % 100 == 99 and judge(G, G.get_residual) < 1e-13:
            break
    rg = judge(G, G.get_residual)
    for it2 in range(4000):
        assert O.setup_preconditioner() == 0
        O.update()
        if it2 % 100 == 99 and judge(O, O.get_res) < 1e-13:
            break
    a, b = G.get_state(), O.get_state()
    dv = a[0] - b[0]
    # the converged state is the uniform stream of the boundary values, whose pressure
    # level is a null mode (ins: the constant pressure; cns: with u, v, w, RT held on every
    # face, a uniform pressure leaves every flux divergence zero): two different
    # iterations agree modulo that constant
    pi = 3 if physics == "ins" else 0
    dv[:, pi] -= dv[:, pi].mean()
    err = max(np.abs(dv).max(), np.abs(a[1] - b[1]).max()) / np.abs(b[0]).max()
    print("outer iterations", it + 1, "final judge", rg, "from", r0, "oracle iterations", it2 + 1,
          "solution rel diff", err, flush=True)
    assert rg < 1e-10 * r0
    assert err < 1e-10


def test_cns_mixed_reaches_fp64_solution(oracle_mod):
    """tri cns FP32 defect correction on a setup where the reference's explicit cns
    relaxation converges (6 x 6 jittered, p = 2, mu = 1e-2, Mach 0.3, amplitude 1e-2,
    CFL 0.05 on every level)."""
    _case("tri cns FP32 defect correction converges to the FP64 solution 6x6 p=2 cfl=0.05")
    from hpmp import _lib, mesh as M, workloads as W
    from hpmp.tri_hp import TriHpCns
    m = M.unit_square(6, 0.2, 3, essential_mask=INFLOW)
    st, bc = W.cns_gaussian_state(m, 2, seed=5, mach=0.3, amp=0.01, width=0.2)
    cfl = (0.05, 0.05, 0.05)
    G = TriHpCns(m, log2p=1, mu=1e-2, bc=bc, cfl=cfl)
    O = oracle_mod.TriHpCns(m, log2p=1, mu=1e-2, bc=bc, cfl=cfl)
    G.set_state(*st)
    O.set_state(*st)
    bv, bs = m.boundary_vertex_mask(), m.boundary_segment_mask()

    def judge(X, get_res):
        X.rsdl()
        v, s, i = [np.array(a) for a in get_res()]
        v[bv, 1:] = 0
        s[bs, :, 1:] = 0
        return max(np.abs(v).max(), np.abs(s).max(), np.abs(i).max() if i.size else 0.0)

    G.set_precision(_lib.HPG_FP32_DEFECT, ninner=1)
    r0 = judge(G, G.get_residual)
    for it in range(3000):
        assert G.setup_preconditioner() == 0
        G.update()
        if it % 100 == 99 and judge(G, G.get_residual) < 1e-12 * r0:
            break
    rg = judge(G, G.get_residual)
    for it2 in range(3000):
        assert O.setup_preconditioner() == 0
        O.update()
        if it2 % 100 == 99 and judge(O, O.get_res) < 1e-12 * r0:
            break
    a, b = G.get_state(), O.get_state()
    # With u, v, RT held on every side the steady cns problem has a family of solutions
    # (nothing fixes the density carried in along each streamline): two FP64 runs of the
    # oracle from states 1e-6 apart converge (judge 5e-16) to states about 2e-8 apart.
    # So the FP32 defect correction is checked by the oracle's own FP64 equations: its
    # residual at the converged mixed state must vanish to the FP64 bar.
    O.set_state(*a)
    ro = judge(O, O.get_res)
    d = [np.array(x) - np.array(y) for x, y in zip(a, b)]
    dist = max(np.abs(x).max() for x in d if x.size) / np.abs(b[0]).max()
    print("outer iterations", it + 1, "final judge", rg, "from", r0, "oracle iterations", it2 + 1,
          "oracle residual at the mixed state", ro, "distance to the oracle's state", dist, flush=True)
    assert rg < 1e-10 * r0
    assert ro < 1e-10 * r0
    assert dist < 1e-5
