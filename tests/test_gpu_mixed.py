"""Mixed-precision path: FP32 residual kernels and FP64 defect correction.

The FP32 residual of the inner solve is pinned against the oracle's FP64 residual at
FP32 accuracy; the exact FP32 Jacobian-vector product (hpg_apply_jacobian, used by the
learned preconditioner's diagonal probe) against central differences of the oracle's
FP64 residual; the defect-corrected solve must reach the FP64 judge's 1e-10 and the
oracle's converged solution to 1e-10.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


def test_mfma_f32_lane_layout():
    _case("mfma f32 16x16x4 layout")
    from hpmp import _lib
    L = _lib.load()
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 4)).astype(np.float32)
    B = rng.standard_normal((4, 16)).astype(np.float32)
    out = np.zeros(256, dtype=np.float32)
    assert L.hpg_selftest_mfma_f32(_lib.ptr(A), _lib.ptr(B), _lib.ptr(out)) == 0
    D = A.astype(np.float64) @ B.astype(np.float64)
    # register i of lane l holds row 4*(l//16) + i (the layout tri_kernels_lin.hip assumes)
    exp = np.array([[D[4 * (l // 16) + i, l % 16] for i in range(4)] for l in range(64)]).ravel()
    alt = np.array([[D[4 * i + l // 16, l % 16] for i in range(4)] for l in range(64)]).ravel()
    print("row 4q+i err", np.abs(out - exp).max(), " row 4i+q err", np.abs(out - alt).max(), flush=True)
    assert np.abs(out - exp).max() < 1e-5


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_jacobian_matches_oracle_central_difference(oracle_mod, log2p):
    _case(f"J*delta vs oracle central difference log2p={log2p}")
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpIns
    p = 1 << log2p
    m = M.unit_square(8, 0.2, 3)
    st = W.noisy_freestream(m, p, seed=21)
    G = TriHpIns(m, log2p=log2p, mu=1e-2)
    G.set_state(*st)
    rng = np.random.default_rng(22)
    d = [rng.standard_normal(a.shape) * 0.1 for a in st]
    jd = G.apply_jacobian(*d)
    O = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    eps = 1e-4
    O.set_state(*[a + eps * b for a, b in zip(st, d)])
    O.rsdl()
    rp = O.get_res()
    O.set_state(*[a - eps * b for a, b in zip(st, d)])
    O.rsdl()
    rm = O.get_res()
    fd = [(a - b) / (2 * eps) for a, b in zip(rp, rm)]
    err = rel_err(jd, fd)
    print("rel err", err, flush=True)
    assert err < 2e-5


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_fp32_residual_matches_oracle(oracle_mod, log2p):
    """The FP32 element and edge kernels of the inner solve (FP32 MFMA layout, operator
    rows permuted back to the FP64 mode order) against the oracle's FP64 residual."""
    _case(f"FP32 residual vs oracle log2p={log2p}")
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpIns
    p = 1 << log2p
    m = M.unit_square(8, 0.2, 3)
    st = W.fourier_state(m, p, seed=23)
    G = TriHpIns(m, log2p=log2p, mu=1e-2)
    G.set_state(*st)
    r32 = G.residual_fp32()
    O = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    O.set_state(*st)
    O.rsdl()
    err = rel_err(r32, O.get_res())
    print("rel err", err, flush=True)
    # FP32 evaluation: rounding of the state and of every product, FP32 accumulation
    assert err < 2e-5
    G.rsdl()
    assert rel_err(G.get_residual(), O.get_res()) < 1e-12


def test_mixed_precision_reaches_fp64_solution(oracle_mod):
    _case("FP32 defect correction converges to the FP64 solution log2p=1 n=6")
    from hpmp import _lib, mesh as M, workloads as W
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(6, 0.2, 15)
    st = W.noisy_freestream(m, 2, seed=16, scale=0.3)
    G = TriHpIns(m, log2p=1, mu=5e-2)
    G.set_state(*st)
    G.rsdl()
    r0 = G.resnorm()
    G.set_precision(_lib.HPG_FP32_DEFECT, ninner=2)
    hist = []
    for it in range(3000):
        G.setup_preconditioner()
        G.update()
        if it % 25 == 24:
            G.rsdl()
            hist.append(G.resnorm())
            if hist[-1] < 1e-11 * r0:
                break
    print("outer iterations", it + 1, "final rel", hist[-1] / r0, flush=True)
    assert hist[-1] < 1e-10 * r0
    O = oracle_mod.TriHpIns(m, log2p=1, mu=5e-2)
    O.set_state(*st)
    for k in range(6000):
        O.setup_preconditioner()
        O.update()
        if k % 200 == 199:
            O.rsdl()
            rv, rs, ri = O.get_res()
            bv = m.boundary_vertex_mask(); bs = m.boundary_segment_mask()
            rv = rv.copy(); rs = rs.copy(); rv[bv, :2] = 0; rs[bs, :, :2] = 0
            if max(np.abs(rv).max(), np.abs(rs).max()) < 1e-12 * r0:
                break
    # a constant pressure is a null mode (velocity prescribed on the whole boundary):
    # different iterations converge to solutions that differ only by it
    from hpmp.tri_hp import rel_diff_mod_pressure
    err = rel_diff_mod_pressure(G.get_state(), O.get_state())
    print("solution rel diff (mod pressure constant)", err, flush=True)
    assert err < 1e-10
