"""Edge cases and error paths of the C-ABI on the GPU: the smallest meshes against the
oracle, index arrays out of range (rejected on the host before any kernel runs),
NaN states (setup_preconditioner reports an error count, as the reference's
`!(qmax >= 0.0)` check does), invalid stages and multigrid levels."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _case(tag):
    print(f"\n[gpu-case] {tag}", flush=True)


@pytest.mark.parametrize("log2p", [0, 1, 2])
def test_single_square_parity(oracle_mod, log2p):
    _case(f"tri ins 1x1 square (2 triangles) log2p={log2p}")
    from hpmp import mesh as M, workloads as W
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(1, 0.0, 0)
    st = W.noisy_freestream(m, 1 << log2p, seed=3, scale=0.3)
    G = TriHpIns(m, log2p=log2p, mu=1e-2)
    O = oracle_mod.TriHpIns(m, log2p=log2p, mu=1e-2)
    for X in (G, O):
        X.set_state(*st)
        X.setup_preconditioner()
        X.update()
        X.rsdl()
    err = rel_err(G.get_residual(), O.get_res())
    print("rel err", err, flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("log2p", [0, 1])
def test_single_cube_parity(oracle_mod, log2p):
    _case(f"tet ins 1 cube (6 tetrahedra) log2p={log2p}")
    import oracle_tet as OT
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpIns
    m = TM.unit_cube(1)
    rng = np.random.default_rng(4)
    v = np.tile([1.0, 0.3, -0.2, 0.5], (m.npnt, 1)) + 0.05 * rng.standard_normal((m.npnt, 4))
    e = 0.05 * rng.standard_normal((m.nseg, (1 << log2p) - 1, 4))
    G = TetHpIns(m, log2p=log2p, mu=1e-2)
    O = OT.TetHpIns(m, log2p=log2p, mu=1e-2)
    for X in (G, O):
        X.set_state(v, e)
        X.setup_preconditioner()
        X.update()
        X.rsdl()
    err = rel_err(G.get_residual(), O.get_res())
    print("rel err", err, flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("log2p", [0, 1])
def test_single_cube_cns_parity(oracle_mod, log2p):
    _case(f"tet cns 1 cube (6 tetrahedra) log2p={log2p}")
    import oracle_tet as OT
    from hpmp import tetmesh as TM
    from hpmp.tet_hp import TetHpCns
    m = TM.unit_cube(1)
    st = np.array([1.0, 0.3, 0.1, -0.05, 1.0])
    rng = np.random.default_rng(6)
    v = np.tile(st, (m.npnt, 1)) + 0.02 * rng.standard_normal((m.npnt, 5))
    e = 0.02 * rng.standard_normal((m.nseg, (1 << log2p) - 1, 5))
    kw = dict(gamma=1.4, prandtl=0.72, R=1.0)
    G = TetHpCns(m, log2p=log2p, mu=1e-2, bc=tuple(st[1:]), **kw)
    O = OT.TetHpCns(m, log2p=log2p, mu=1e-2, bc=(0.0, *st[1:]), **kw)
    for X in (G, O):
        X.set_state(v, e)
        assert X.setup_preconditioner() == 0
        X.update()
        X.rsdl()
    err = rel_err(G.get_residual(), O.get_res())
    print("rel err", err, flush=True)
    assert err < 1e-12


def test_out_of_range_indices_are_rejected():
    _case("index arrays out of range")
    from hpmp import _lib, mesh as M, tetmesh as TM
    from hpmp.tet_hp import TetHpIns
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(2, 0.0, 0)
    m.tri_pnt = m.tri_pnt.copy()
    m.tri_pnt[3, 1] = m.npnt          # one past the end
    with pytest.raises(_lib.HpgError):
        TriHpIns(m, log2p=1)
    m = M.unit_square(2, 0.0, 0)
    m.tri_sgn = m.tri_sgn.copy()
    m.tri_sgn[0, 0] = 0
    with pytest.raises(_lib.HpgError):
        TriHpIns(m, log2p=1)
    t = TM.unit_cube(2)
    t.tet_seg = t.tet_seg.copy()
    t.tet_seg[5, 2] = -1
    with pytest.raises(_lib.HpgError):
        TetHpIns(t, log2p=1)


def test_mesh_without_edge_boundaries_and_null_arrays():
    """nebd = 0 with NULL boundary arrays is a valid mesh (no edge boundaries); a
    non-empty boundary with a NULL segment list is rejected on the host."""
    _case("nebd = 0 with null boundary arrays")
    import ctypes as C
    from hpmp import _lib, mesh as M
    L = _lib.load()
    m = M.unit_square(2, 0.0, 0)
    k = [np.ascontiguousarray(a) for a in (m.pnts.astype(np.float64), m.seg_pnt.astype(np.int32),
                                           m.tri_pnt.astype(np.int32), m.tri_seg.astype(np.int32),
                                           m.tri_sgn.astype(np.int32))]
    prm = _lib.ParamsC()
    prm.physics, prm.log2p, prm.rho, prm.mu, prm.adis, prm.fadd, prm.nstage = 0, 1, 1.0, 1e-2, 1.0, 1.0, 1
    prm.alpha[0] = prm.beta[0] = 1.0
    prm.cfl[0] = prm.cfl[1] = 1.0
    ptrs = [C.c_void_p(a.ctypes.data) for a in k]
    cm = _lib.TriMeshC(m.npnt, m.nseg, m.ntri, *ptrs, 0, None, None, None)
    h = C.c_void_p()
    assert L.hpg_create(C.byref(cm), C.byref(prm), C.byref(h)) == 0 and h.value
    assert L.hpg_rsdl(h, 1) == 0
    assert L.hpg_destroy(h) == 0
    start = np.array([0, 2], dtype=np.int32)
    ess = np.array([3], dtype=np.int32)
    cm = _lib.TriMeshC(m.npnt, m.nseg, m.ntri, *ptrs, 1, C.c_void_p(start.ctypes.data), None,
                       C.c_void_p(ess.ctypes.data))
    h = C.c_void_p()
    assert L.hpg_create(C.byref(cm), C.byref(prm), C.byref(h)) < 0


def test_nan_state_reports_setup_error(oracle_mod):
    _case("NaN state -> setup_preconditioner error count")
    from hpmp import mesh as M, tetmesh as TM, workloads as W
    from hpmp.tet_hp import TetHpIns
    from hpmp.tri_hp import TriHpIns
    m = M.unit_square(4, 0.1, 1)
    v, s, i = W.noisy_freestream(m, 4, seed=1)
    v = v.copy()
    v[7, 0] = np.nan
    G = TriHpIns(m, log2p=2, mu=1e-2)
    G.set_state(v, s, i)
    assert G.setup_preconditioner() >= 1
    # tet: qmax = MAX(qmax, q) over the element's vertices keeps a NaN only from the
    # last vertex, so the error count follows the vertex position; compare with the
    # oracle vertex by vertex
    import oracle_tet as OT
    t = TM.unit_cube(2)
    v0 = np.tile([1.0, 0.0, 0.0, 0.0], (t.npnt, 1))
    e0 = np.zeros((t.nseg, 1, 4))
    T = TetHpIns(t, log2p=1, mu=1e-2)
    O = OT.TetHpIns(t, log2p=1, mu=1e-2)
    seen = set()
    for pos in range(t.npnt):
        vt = v0.copy()
        vt[pos, pos % 3] = np.nan
        T.set_state(vt, e0)
        O.set_state(vt, e0)
        ge, oe = T.setup_preconditioner(), O.setup_preconditioner()
        assert (ge > 0) == (oe > 0), pos
        seen.add(oe)
    assert seen == {0, 1}
    T.set_state(v0, e0)
    assert T.setup_preconditioner() == 0


def test_invalid_stage_and_levels():
    _case("invalid stage and multigrid levels")
    from hpmp import _lib, mesh as M, tetmesh as TM
    from hpmp.tet_hp import TetHpIns
    from hpmp.tri_hp import TriHpIns
    G = TriHpIns(M.unit_square(2, 0.0, 0), log2p=1)
    with pytest.raises(_lib.HpgError):
        G.rsdl(99)
    with pytest.raises(_lib.HpgError):
        G.mg_prolongate()            # already at the finest level
    G.mg_restrict()
    with pytest.raises(_lib.HpgError):
        G.mg_restrict()              # no level below p = 1
    T = TetHpIns(TM.unit_cube(2), log2p=0)
    with pytest.raises(_lib.HpgError):
        T.mg_restrict()
    with pytest.raises(_lib.HpgError):
        T.rsdl(-1)
