"""tri_hp_ins on the GPU: the reference's per-block interface over libhpgpu.so.

``TriHpIns`` keeps the method names and meanings of tri_hp / tri_hp_ins
(tri_hp/tri_hp.h:207-252, tri_hp/ins/tri_hp_ins.h:43-46): ``rsdl(stage)``,
``setup_preconditioner() -> int``, ``update()``, ``minvrt()``, ``mg_restrict()``,
``mg_prolongate()``, ``maxres()``; ``cycle`` restates block::cycle
(tri_mesh/blocks.cpp:1267-1348) for one grid with ``extra_finest_levels``
p-levels.  State lives on the device between ``set_state`` and ``get_state``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .input_map import InputMap
from .mesh import TriMesh

NSTAGE_ALPHA = (0.25, 1.0 / 6.0, 0.375, 0.5, 1.0)   # tri_mesh/blocks.cpp:1133
NSTAGE_BETA = (1.0, 0.0, 5.0 / 9.0, 0.0, 4.0 / 9.0)  # tri_mesh/blocks.cpp:1137
CFL_DEFAULT = (2.5, 1.5, 1.0, 0.5)                   # tri_hp/allocate.cpp:275


def rel_diff_mod_pressure(a, b) -> float:
    """max |a - b| / max |b| over (v, s, i) after removing the constant-pressure offset
    (mean over vertices of the pressure difference).  With the velocity prescribed on
    the whole boundary a constant pressure is a null mode of the ins residual: the
    Dirichlet rows drop the boundary momentum equations, interior test functions
    integrate a constant pressure to zero and the edge flux uses the boundary
    pressure, so converged solutions of two different iterations agree only up to it.
    The constant is spanned by the vertex modes alone in the hierarchical basis."""
    d = [np.array(x, dtype=np.float64) - np.asarray(y, dtype=np.float64) for x, y in zip(a, b)]
    d[0][:, 2] -= d[0][:, 2].mean()
    num = max(float(np.abs(x).max()) for x in d if x.size)
    den = max(float(np.abs(np.asarray(y)).max()) for y in b if np.asarray(y).size)
    return num / den if den > 0 else num


class TriHpIns:
    NV = 3
    PHYSICS = _lib.HPG_PHYSICS_INS

    def __init__(self, mesh: TriMesh, log2p: int = 2, rho: float = 1.0, mu: float = 0.0, adis: float = 1.0,
                 bc_pressure: float = 0.0, fadd: float = 1.0, alpha=NSTAGE_ALPHA, beta=NSTAGE_BETA,
                 cfl=CFL_DEFAULT, **extra):
        self.L = _lib.load()
        self.mesh = mesh
        self.log2pmax = int(log2p)
        self.p = 1 << self.log2pmax
        self.sm0 = self.p - 1
        self.im0 = (self.p - 1) * (self.p - 2) // 2
        self.nstage = len(alpha)
        m = mesh
        self._keep = [np.ascontiguousarray(a) for a in (
            m.pnts.astype(np.float64), m.seg_pnt.astype(np.int32), m.tri_pnt.astype(np.int32),
            m.tri_seg.astype(np.int32), m.tri_sgn.astype(np.int32), m.ebdry_start.astype(np.int32),
            m.ebdry_seg.astype(np.int32), m.ebdry_essential.astype(np.int32))]
        k = self._keep
        cm = _lib.TriMeshC(m.npnt, m.nseg, m.ntri, *(C.c_void_p(a.ctypes.data) for a in k[:5]), m.nebd,
                           *(C.c_void_p(a.ctypes.data) for a in k[5:]))
        prm = _lib.ParamsC()
        prm.physics = self.PHYSICS
        prm.log2p = self.log2pmax
        prm.rho, prm.mu, prm.adis, prm.bc_pressure, prm.fadd = rho, mu, adis, bc_pressure, fadd
        prm.nstage = self.nstage
        for s in range(self.nstage):
            prm.alpha[s] = alpha[s]
            prm.beta[s] = beta[s]
        for lvl in range(4):
            prm.cfl[lvl] = cfl[lvl] if lvl < len(cfl) else 0.5
        self.params = dict(rho=rho, mu=mu, adis=adis, bc_pressure=bc_pressure, fadd=fadd,
                           alpha=tuple(alpha), beta=tuple(beta), cfl=tuple(cfl), **extra)
        self._fill_extra(prm, **extra)
        h = C.c_void_p()
        rc = self.L.hpg_create(C.byref(cm), C.byref(prm), C.byref(h))
        if rc != 0 or not h.value:
            raise _lib.HpgError(f"hpg_create failed ({rc})")
        self.h = h

    def _fill_extra(self, prm, **extra):
        if extra:
            raise TypeError(f"unexpected arguments {sorted(extra)}")

    @classmethod
    def from_input_map(cls, mesh: TriMesh, inmap: InputMap, prefix: str = "b0") -> "TriHpIns":
        log2p = inmap.geti("log2p", 0, prefix)
        nstage = inmap.geti("nstage", 5, prefix)
        alpha = inmap.getlist("alpha", NSTAGE_ALPHA, nstage, prefix)
        beta = inmap.getlist("beta", NSTAGE_BETA, nstage, prefix)
        cfl = inmap.getlist("cfl", CFL_DEFAULT, log2p + 1, prefix)
        return cls(mesh, log2p=log2p, rho=inmap.getf("rho", 1.0, prefix), mu=inmap.getf("mu", 0.0, prefix),
                   adis=inmap.getf("dissipation", 1.0, prefix), fadd=inmap.getf("hp_fadd", 1.0, prefix),
                   alpha=alpha, beta=beta, cfl=list(cfl) + list(CFL_DEFAULT[len(cfl):]))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.hpg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = self.L.hpg_last_error(self.h).decode()
            raise _lib.HpgError(f"{what} failed ({rc}): {msg}")
        return rc

    # --- array exchange (reference layouts) ------------------------------------------
    def shapes(self):
        m = self.mesh
        return ((m.npnt, self.NV), (m.nseg, self.sm0, self.NV), (m.ntri, self.im0, self.NV))

    def zeros(self):
        return tuple(np.zeros(s) for s in self.shapes())

    def set_state(self, v, s, i):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (v, s, i)]
        self._check(self.L.hpg_set_state(self.h, *(_lib.ptr(x) for x in a)), "set_state")

    def get_state(self):
        out = self.zeros()
        self._check(self.L.hpg_get_state(self.h, *(_lib.ptr(x) for x in out)), "get_state")
        return out

    def get_residual(self):
        out = self.zeros()
        self._check(self.L.hpg_get_residual(self.h, *(_lib.ptr(x) for x in out)), "get_residual")
        return out

    def set_residual(self, v, s, i):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (v, s, i)]
        self._check(self.L.hpg_set_residual(self.h, *(_lib.ptr(x) for x in a)), "set_residual")

    def get_preconditioner(self):
        m = self.mesh
        vp, sp, tp = np.zeros((m.npnt, self.NV)), np.zeros((m.nseg, self.NV)), np.zeros((m.ntri, self.NV))
        self._check(self.L.hpg_get_preconditioner(self.h, _lib.ptr(vp), _lib.ptr(sp), _lib.ptr(tp)), "get_prcn")
        return vp, sp, tp

    def set_tprcn(self, tp):
        tp = np.ascontiguousarray(tp, dtype=np.float64)
        self._check(self.L.hpg_set_tprcn(self.h, _lib.ptr(tp)), "set_tprcn")

    # --- the reference's virtual interface ----------------------------------------------
    def rsdl(self, stage: int | None = None):
        self._check(self.L.hpg_rsdl(self.h, self.nstage if stage is None else int(stage)), "rsdl")

    def setup_preconditioner(self) -> int:
        return self._check(self.L.hpg_setup_preconditioner(self.h), "setup_preconditioner")

    def minvrt(self):
        self._check(self.L.hpg_minvrt(self.h), "minvrt")

    def update(self):
        self._check(self.L.hpg_update(self.h), "update")

    def maxres(self):
        pv = np.zeros(self.NV)
        e = np.zeros(1)
        self._check(self.L.hpg_maxres(self.h, _lib.ptr(pv), _lib.ptr(e)), "maxres")
        return float(e[0]), pv

    def resnorm(self) -> float:
        e = np.zeros(1)
        self._check(self.L.hpg_resnorm(self.h, _lib.ptr(e)), "resnorm")
        return float(e[0])

    def mg_restrict(self):
        self._check(self.L.hpg_mg_restrict(self.h), "mg_restrict")

    def mg_prolongate(self):
        self._check(self.L.hpg_mg_prolongate(self.h), "mg_prolongate")

    @property
    def log2p(self) -> int:
        v = C.c_int()
        self._check(self.L.hpg_get_log2p(self.h, C.byref(v)), "get_log2p")
        return v.value

    def to_finest(self):
        """Return to the finest p-level (after a cycle was abandoned on a coarse one)."""
        while self.log2p < self.log2pmax:
            self.mg_prolongate()

    def set_precision(self, mode: int, ninner: int | None = None):
        """HPG_FP64 (plain) or HPG_FP32_DEFECT (FP64 defect correction around `ninner`
        multistage iterations with the FP32 element residual), ins and cns."""
        self._check(self.L.hpg_set_precision(self.h, mode), "set_precision")
        if ninner is not None:
            self._check(self.L.hpg_set_defect_inner(self.h, int(ninner)), "set_defect_inner")

    def set_cycle_graph(self, enable: bool):
        """Graph capture of mg_cycle on this (fine) context: on by default."""
        self._check(self.L.hpg_set_cycle_graph(self.h, int(bool(enable))), "set_cycle_graph")

    def set_cfl(self, cfl):
        """Pseudo-time CFL per p-level (the reference's per-level ``cfl`` input)."""
        a = np.ascontiguousarray(cfl, dtype=np.float64)
        self._check(self.L.hpg_set_cfl(self.h, a.ctypes.data_as(C.POINTER(C.c_double)), int(a.size)), "set_cfl")

    def set_periodic(self, pairs):
        """Periodic boundaries (ins): ``pairs`` of edge boundaries (b1, b2), side j of b1
        matched with side n-1-j of b2 (the reference's prdc pairing).  Call before
        set_edge_flux and set_curved; makes the state's copies equal."""
        pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        self._check(self.L.hpg_set_periodic(self.h, int(pr.shape[0]), pr.ctypes.data_as(C.POINTER(C.c_int))),
                    "set_periodic")

    def set_edge_flux(self, types, ibc=(0.0, 0.0, 0.0)):
        """Boundary flux of each edge boundary (ins): HPG_EDGE_GENERIC (inflow, outflow,
        symmetry), HPG_EDGE_EULER or HPG_EDGE_CHARACTERISTIC, and the freestream state
        ``ibc`` = (u, v, p) of the initial/boundary-condition function."""
        t = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        b = np.ascontiguousarray(ibc, dtype=np.float64).reshape(3)
        self._check(self.L.hpg_set_edge_flux(self.h, int(t.size), t.ctypes.data_as(C.POINTER(C.c_int)),
                                             b.ctypes.data_as(C.POINTER(C.c_double))), "set_edge_flux")

    def set_curved(self, sides, crds):
        """Curved boundary sides (ins): ``crds`` (nside, sm0, 2), the coordinate side modes
        of each listed boundary side in its own direction, as the reference's
        ``hp_edge_bdry::crds`` (hpmp.curved.side_modes computes them from a boundary
        shape).  An empty list makes every side straight again."""
        sd = np.ascontiguousarray(sides, dtype=np.int32).reshape(-1)
        cr = np.ascontiguousarray(crds, dtype=np.float64).reshape(sd.size, self.sm0, 2)
        self._check(self.L.hpg_set_curved(self.h, int(sd.size), sd.ctypes.data_as(C.POINTER(C.c_int)),
                                          cr.ctypes.data_as(C.POINTER(C.c_double))), "set_curved")

    def set_preconditioner_mode(self, mode: int):
        """HPG_PRCN_ANALYTIC or HPG_PRCN_LEARNED (needs an MLP attached, hpmp.learned)."""
        self._check(self.L.hpg_set_preconditioner(self.h, mode), "set_preconditioner")

    def apply_jacobian(self, dv, ds, di):
        """J(u) delta at the current state (FP32 linearised kernels); leaves R(u) in the residual."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (dv, ds, di)]
        out = self.zeros()
        self._check(self.L.hpg_apply_jacobian(self.h, *(_lib.ptr(x) for x in a), *(_lib.ptr(x) for x in out)),
                    "apply_jacobian")
        return out

    def mg_cycle(self, grids, extra_finest_levels: int, vw: int = 1, itercrsn: int = 1, iterrfne: int = 0) -> float:
        """One block::cycle over `grids` (this context first) in one library call
        (hpg_mg_cycle); returns the entry level's maxres; FloatingPointError as
        hpmp.multigrid.cycle."""
        hs = (C.c_void_p * len(grids))(*[g.h.value for g in grids])
        mx = C.c_double()
        rc = self.L.hpg_mg_cycle(hs, len(grids), int(extra_finest_levels), int(vw), int(itercrsn), int(iterrfne),
                                 C.byref(mx))
        self._check(rc, "mg_cycle")
        if rc > 0:
            raise FloatingPointError("setup_preconditioner reported an error")
        if not mx.value > 0.0:
            raise FloatingPointError("maxres is not positive (NaN or exact zero)")
        return float(mx.value)

    def residual_fp32(self):
        """The full residual evaluated in FP32 by the defect correction's inner-solve kernels."""
        out = self.zeros()
        self._check(self.L.hpg_residual_fp32(self.h, *(_lib.ptr(x) for x in out)), "residual_fp32")
        return out

    def element_jacobian(self, tinds, eps_r: float = 1e-8, eps_a: float = 1e-8):
        """tri_hp::element_jacobian (tri_hp/nstage.cpp:131-165) for the distinct elements
        ``tinds`` on the current level: array (len, N, N), N = tm * NV, with
        K[i*NV + n, mode*NV + var] = d(lf_re + lf_im)(n)(i) / d uht(var)(mode) by forward
        differences (the reference's eps_r = eps_a = 1e-8)."""
        t = np.ascontiguousarray(np.atleast_1d(tinds), dtype=np.int32)
        p = 1 << self.log2p
        n = (p + 1) * (p + 2) // 2 * self.NV
        K = np.zeros((t.size, n, n))
        self._check(self.L.hpg_element_jacobian(self.h, t.ctypes.data_as(C.POINTER(C.c_int)), int(t.size),
                                                float(eps_r), float(eps_a), _lib.ptr(K)), "element_jacobian")
        return K

    def synchronize(self):
        self._check(self.L.hpg_synchronize(self.h), "synchronize")

    def link_fine(self, fine: "TriHpIns", cc_tri, cc_wt, fc_tri, fc_wt):
        """Make this level (a coarser mesh at log2p = 0) the h-multigrid neighbour of
        ``fine`` (tri_mesh::connect, tri_mesh/blocks.cpp:1245-1260); afterwards
        mg_restrict/mg_prolongate on this level move between the two meshes
        (tri_hp/movco.cpp:41-102, tri_hp/movfin.cpp:16-82)."""
        m, f = self.mesh, fine.mesh
        cc_tri = np.ascontiguousarray(cc_tri, dtype=np.int32)
        fc_tri = np.ascontiguousarray(fc_tri, dtype=np.int32)
        cc_wt = np.ascontiguousarray(cc_wt, dtype=np.float64)
        fc_wt = np.ascontiguousarray(fc_wt, dtype=np.float64)
        if cc_tri.shape != (f.npnt,) or cc_wt.shape != (f.npnt, 3):
            raise ValueError("cc_tri/cc_wt must hold one coarse triangle and 3 weights per fine vertex")
        if fc_tri.shape != (m.npnt,) or fc_wt.shape != (m.npnt, 3):
            raise ValueError("fc_tri/fc_wt must hold one fine triangle and 3 weights per coarse vertex")
        if cc_tri.min() < 0 or cc_tri.max() >= m.ntri or fc_tri.min() < 0 or fc_tri.max() >= f.ntri:
            raise ValueError("connection triangle index out of range")
        self._check(self.L.hpg_link_fine(self.h, fine.h, _lib.ptr(cc_tri), _lib.ptr(cc_wt), _lib.ptr(fc_tri),
                                         _lib.ptr(fc_wt)), "link_fine")
        self._fine = fine

    def cycle(self, vw: int = 1, mglvls: int = 1, itercrsn: int = 1, iterrfne: int = 0, lvl: int = 0) -> float:
        """block::cycle (tri_mesh/blocks.cpp:1267-1348) on one grid; p-levels below."""
        error = 0.0
        for vcount in range(vw):
            if self.setup_preconditioner():
                raise FloatingPointError("setup_preconditioner reported an error")
            for _ in range(itercrsn):
                self.update()
            if lvl + vcount == 0:
                error, _ = self.maxres()
                if not error > 0.0:
                    raise FloatingPointError("maxres is not positive (NaN or exact zero)")
            if lvl == mglvls - 1:
                return error
            self.rsdl()
            self.mg_restrict()
            self.cycle(vw, mglvls, itercrsn, iterrfne, lvl + 1)
            self.mg_prolongate()
            if iterrfne:
                if self.setup_preconditioner():
                    raise FloatingPointError("setup_preconditioner reported an error")
            for _ in range(iterrfne):
                self.update()
        return error


CNS_CFL_DEFAULT = (2.5, 1.5, 1.0)


class TriHpCns(TriHpIns):
    """tri_hp_cns on the GPU (tri_hp/cns/tri_hp_cns.h:19-72): primitive unknowns
    (p, u, v, RT), NV = 4.  Boundaries with essential mask 0b1110 are inflow
    boundaries at the constant state `bc` = (p, u, v, RT), which also gives the
    boundary pressure of the edge flux (tri_hp/cns/bdry.cpp:192-245)."""

    NV = 4
    PHYSICS = _lib.HPG_PHYSICS_CNS

    def __init__(self, mesh: TriMesh, log2p: int = 2, mu: float = 1e-3, adis: float = 1.0, gamma: float = 1.4,
                 prandtl: float = 0.713, Rgas: float = 1.0, bc=(1.0, 0.5, 0.0, 1.0), fadd: float = 1.0,
                 alpha=NSTAGE_ALPHA, beta=NSTAGE_BETA, cfl=CNS_CFL_DEFAULT):
        super().__init__(mesh, log2p=log2p, rho=1.0, mu=mu, adis=adis, bc_pressure=bc[0], fadd=fadd, alpha=alpha,
                         beta=beta, cfl=cfl, gamma=gamma, prandtl=prandtl, Rgas=Rgas, bc=tuple(bc))

    def _fill_extra(self, prm, gamma, prandtl, Rgas, bc):
        prm.gamma, prm.prandtl, prm.Rgas = gamma, prandtl, Rgas
        for n in range(4):
            prm.bc[n] = bc[n]

    def get_block_preconditioner(self):
        m = self.mesh
        out = [np.zeros((m.npnt, 4, 4)), np.zeros((m.nseg, 4, 4)), np.zeros((m.ntri, 4, 4))]
        self._check(self.L.hpg_get_block_preconditioner(self.h, *(_lib.ptr(x) for x in out)), "get_block_preconditioner")
        return out
