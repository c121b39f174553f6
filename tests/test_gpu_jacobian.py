"""GPU element Jacobian (hpg_element_jacobian, tri_hp/nstage.cpp:131-165) against the
oracle's tri_hp::element_jacobian, ins and cns.

Forward differences amplify the rounding difference of the two residual evaluations
by 1/dw: with residuals that agree to about 1e-14 of their size and dw near 3e-8 (the
reference's eps_r = eps_a = 1e-8), the columns agree to about 1e-6 of max |K|.  With
eps_r = eps_a = 1e-4 the amplification drops to about 1e-10 while the difference
quotient itself changes by O(dw) against the 1e-8 one, so a dw formula or a column
convention that differs from the oracle's shows up far above that tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
INFLOW = 0b1110


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def _ins_pair(oracle_mod, log2p, n=5, seed=2):
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(n, 0.2, seed)
    st = W.noisy_freestream(m, 1 << log2p, seed=seed + 5, scale=0.3)
    G = TriHpIns(m, log2p=log2p, mu=1e-2)
    O = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    G.set_state(*st)
    O.set_state(*st)
    return m, G, O


def _cns_pair(oracle_mod, log2p, n=5, seed=3):
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpCns
    m = M.unit_square(n, 0.2, seed, essential_mask=INFLOW)
    st, bc = W.cns_gaussian_state(m, 1 << log2p, seed=seed + 8, amp=0.05, width=0.2, noise=2e-2)
    G = TriHpCns(m, log2p=log2p, mu=1e-2, bc=bc)
    O = oracle_mod.TriHpCns(m, log2p=log2p, mu=1e-2, bc=bc)
    G.set_state(*st)
    O.set_state(*st)
    return m, G, O


def _batch(m, k=9, seed=0):
    # distinct elements from every colour, in scrambled order
    rng = np.random.default_rng(seed)
    return rng.choice(m.ntri, size=min(k, m.ntri), replace=False)


def _compare(G, O, tinds, eps, tol):
    K = G.element_jacobian(tinds, *eps)
    worst = 0.0
    for j, t in enumerate(tinds):
        Ko = O.element_jacobian(int(t), *eps)
        assert K[j].shape == Ko.shape
        err = np.abs(K[j] - Ko).max() / np.abs(Ko).max()
        worst = max(worst, err)
    print("max rel err", worst, flush=True)
    assert worst < tol


@pytest.mark.parametrize("log2p", [0, 1, 2])
@pytest.mark.parametrize("eps,tol", [((1e-8, 1e-8), 1e-5), ((1e-4, 1e-4), 1e-9)])
def test_ins_element_jacobian_parity(oracle_mod, log2p, eps, tol):
    _case(f"ins element_jacobian log2p={log2p} eps={eps}")
    m, G, O = _ins_pair(oracle_mod, log2p)
    _compare(G, O, _batch(m, 11, log2p), eps, tol)


@pytest.mark.parametrize("log2p", [0, 1, 2])
@pytest.mark.parametrize("eps,tol", [((1e-8, 1e-8), 1e-5), ((1e-4, 1e-4), 1e-9)])
def test_cns_element_jacobian_parity(oracle_mod, log2p, eps, tol):
    _case(f"cns element_jacobian log2p={log2p} eps={eps}")
    m, G, O = _cns_pair(oracle_mod, log2p)
    _compare(G, O, _batch(m, 9, log2p + 3), eps, tol)


def test_element_jacobian_leaves_state_and_residual_untouched(oracle_mod):
    _case("ins element_jacobian state/residual untouched, whole mesh")
    m, G, O = _ins_pair(oracle_mod, 2)
    G.rsdl()
    s0 = [a.copy() for a in G.get_state()]
    r0 = [a.copy() for a in G.get_residual()]
    K = G.element_jacobian(np.arange(m.ntri))
    assert np.isfinite(K).all()
    for a, b in zip(G.get_state(), s0):
        assert np.array_equal(a, b)
    for a, b in zip(G.get_residual(), r0):
        assert np.array_equal(a, b)
    # the whole-mesh batch equals one-element calls
    for t in (0, m.ntri // 2, m.ntri - 1):
        assert np.array_equal(G.element_jacobian([t])[0], K[t])


def test_element_jacobian_diagonal_matches_linearised_probe(oracle_mod):
    """sum over modes of |K(mn, mn)| against hpg_measure_jacobian_diag, the exact
    derivative through the FP32 linearised kernel: two independent GPU paths."""
    _case("ins element_jacobian diagonal vs FP32 linearised probe")
    from hpmp import learned
    m, G, O = _ins_pair(oracle_mod, 2, n=6)
    K = G.element_jacobian(np.arange(m.ntri))
    tm = K.shape[1] // 3
    p = 4   # log2p = 2
    assert tm == (p + 1) * (p + 2) // 2
    dg = np.abs(np.diagonal(K, axis1=1, axis2=2)).reshape(m.ntri, tm, 3)
    # per mode type: vertex modes 0-2, side modes 3 .. 3p-1, interior modes 3p ..
    dK = np.concatenate([dg[:, :3].sum(axis=1), dg[:, 3:3 * p].sum(axis=1), dg[:, 3 * p:].sum(axis=1)], axis=1)
    D = learned.jacobian_diag(G)
    assert D.shape == (m.ntri, 9)
    err = np.abs(dK - D).max() / np.abs(D).max()
    print("rel diff", err, flush=True)
    assert err < 1e-4


def test_element_jacobian_rejects_bad_batches(oracle_mod):
    _case("element_jacobian rejects out-of-range and repeated elements")
    m, G, O = _ins_pair(oracle_mod, 1, n=2)
    for bad in ([0, m.ntri], [-1], [3, 1, 3]):
        with pytest.raises(RuntimeError):
            G.element_jacobian(bad)
    assert G.element_jacobian([]).shape == (0, 18, 18)
