"""Host-side mirror of tet_hp / tet_hp_ins / tet_hp_cns (tet_hp/tet_hp.h:229-273,
tet_hp/ins/tet_hp_ins.h, tet_hp/cns/tet_hp_cns.h) over the hpg_tet_* C-ABI of
libhpgpu.so.

Same method names and argument meaning as the reference's virtual interface:
rsdl(stage), setup_preconditioner() -> error count, minvrt(), update(), maxres(),
mg_restrict(), mg_prolongate().  State is exchanged in the reference layouts
v(npnt, NV) and e(nseg, em0, NV) (tet_hp/tet_hp.h:89-94).  There is no CPU path:
the library must be built.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .tetmesh import TetMesh

NSTAGE_ALPHA = (0.25, 1.0 / 6.0, 0.375, 0.5, 1.0)
NSTAGE_BETA = (1.0, 0.0, 5.0 / 9.0, 0.0, 4.0 / 9.0)
TET_CFL_DEFAULT = (2.2, 1.5, 1.0, 0.5)   # tet_hp/allocate.cpp:234


class _TetHp:
    NV = 4

    def _create(self, mesh: TetMesh, prm, log2p: int):
        self.L = _lib.load()
        self.mesh = mesh
        self.log2pmax = int(log2p)
        self.em0 = (1 << self.log2pmax) - 1
        m = mesh
        self._keep = [np.ascontiguousarray(a) for a in (
            m.pnts.astype(np.float64), m.seg_pnt.astype(np.int32), m.tet_pnt.astype(np.int32),
            m.tet_seg.astype(np.int32), m.tet_sgn.astype(np.int32), m.bf_pnt.astype(np.int32),
            m.bf_seg.astype(np.int32), m.bf_sgn.astype(np.int32))]
        cm = _lib.TetMeshC(m.npnt, m.nseg, m.ntet, m.nbf, *(C.c_void_p(a.ctypes.data) for a in self._keep))
        h = C.c_void_p()
        rc = self.L.hpg_tet_create(C.byref(cm), C.byref(prm), C.byref(h))
        if rc != 0 or not h.value:
            raise _lib.HpgError(f"hpg_tet_create failed ({rc})")
        self.h = h

    @staticmethod
    def _stages(prm, alpha, beta, cfl):
        prm.nstage = len(alpha)
        for s in range(len(alpha)):
            prm.alpha[s] = alpha[s]
            prm.beta[s] = beta[s]
        for lvl in range(4):
            prm.cfl[lvl] = cfl[lvl] if lvl < len(cfl) else 0.5

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.hpg_tet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> int:
        if rc < 0:
            raise _lib.HpgError(f"{what} failed ({rc}): {self.L.hpg_tet_last_error(self.h).decode()}")
        return rc

    def shapes(self):
        return (self.mesh.npnt, self.NV), (self.mesh.nseg, self.em0, self.NV)

    def zeros(self):
        return tuple(np.zeros(s) for s in self.shapes())

    def set_state(self, v, e):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (v, e)]
        self._check(self.L.hpg_tet_set_state(self.h, *(_lib.ptr(x) for x in a)), "set_state")

    def get_state(self):
        out = self.zeros()
        self._check(self.L.hpg_tet_get_state(self.h, *(_lib.ptr(x) for x in out)), "get_state")
        return out

    def get_residual(self):
        out = self.zeros()
        self._check(self.L.hpg_tet_get_residual(self.h, *(_lib.ptr(x) for x in out)), "get_residual")
        return out

    def set_residual(self, v, e):
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (v, e)]
        self._check(self.L.hpg_tet_set_residual(self.h, *(_lib.ptr(x) for x in a)), "set_residual")

    def get_preconditioner(self):
        """vprcn, eprcn, iprcn (NV values each) and tau: (ntet, 4) for ins, (ntet, 5, 5) for cns"""
        m, nv = self.mesh, self.NV
        out = (np.zeros((m.npnt, nv)), np.zeros((m.nseg, nv)), np.zeros((m.ntet, nv)),
               np.zeros((m.ntet, 4)) if nv == 4 else np.zeros((m.ntet, 5, 5)))
        self._check(self.L.hpg_tet_get_preconditioner(self.h, *(_lib.ptr(x) for x in out)), "get_preconditioner")
        return out

    def rsdl(self, stage: int | None = None):
        self._check(self.L.hpg_tet_rsdl(self.h, self.nstage if stage is None else int(stage)), "rsdl")

    def setup_preconditioner(self) -> int:
        return self._check(self.L.hpg_tet_setup_preconditioner(self.h), "setup_preconditioner")

    def minvrt(self):
        self._check(self.L.hpg_tet_minvrt(self.h), "minvrt")

    def update(self):
        self._check(self.L.hpg_tet_update(self.h), "update")

    def maxres(self):
        pv = np.zeros(self.NV)
        e = np.zeros(1)
        self._check(self.L.hpg_tet_maxres(self.h, _lib.ptr(pv), _lib.ptr(e)), "maxres")
        return float(e[0]), pv

    def resnorm(self) -> float:
        e = np.zeros(1)
        self._check(self.L.hpg_tet_resnorm(self.h, _lib.ptr(e)), "resnorm")
        return float(e[0])

    def mg_restrict(self):
        self._check(self.L.hpg_tet_mg_restrict(self.h), "mg_restrict")

    def mg_prolongate(self):
        self._check(self.L.hpg_tet_mg_prolongate(self.h), "mg_prolongate")

    @property
    def log2p(self) -> int:
        v = C.c_int()
        self._check(self.L.hpg_tet_get_log2p(self.h, C.byref(v)), "get_log2p")
        return v.value

    def element_jacobian(self, tinds, eps_r: float = 1e-6, eps_a: float = 1e-10):
        """tet_hp::element_jacobian (tet_hp/update.cpp:134-178) for the distinct
        tetrahedra ``tinds`` on the current level: array (len, N, N), N = tm * NV, with
        K[i*NV + n, mode*NV + var] by forward differences (one dw for all variables,
        the reference's eps_r = 1e-6, eps_a = 1e-10).  Needs setup_preconditioner first
        (the element residual reads its tau)."""
        t = np.ascontiguousarray(np.atleast_1d(tinds), dtype=np.int32)
        n = (4 if self.log2p == 0 else 10) * self.NV
        K = np.zeros((t.size, n, n))
        self._check(self.L.hpg_tet_element_jacobian(self.h, t.ctypes.data_as(C.POINTER(C.c_int)), int(t.size),
                                                    float(eps_r), float(eps_a), _lib.ptr(K)), "element_jacobian")
        return K

    def set_precision(self, mode: int, ninner: int = 1):
        """HPG_FP64 (plain) or HPG_FP32_DEFECT: FP64 defect correction around `ninner`
        multistage iterations whose element residuals run in FP32 arithmetic."""
        self._check(self.L.hpg_tet_set_precision(self.h, int(mode), int(ninner)), "tet_set_precision")

    def residual_fp32(self):
        """The full residual with the element residual in FP32 arithmetic (inner-solve kernels)."""
        v, e = self.zeros()
        self._check(self.L.hpg_tet_residual_fp32(self.h, _lib.ptr(v), _lib.ptr(e)), "tet_residual_fp32")
        return v, e

    def synchronize(self):
        self._check(self.L.hpg_tet_synchronize(self.h), "synchronize")

    def cycle(self, vw: int = 1, mglvls: int = 1, itercrsn: int = 1, iterrfne: int = 0, lvl: int = 0) -> float:
        """block::cycle (tri_mesh/blocks.cpp:1267-1348) over the p-levels of this block."""
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
            if iterrfne and self.setup_preconditioner():
                raise FloatingPointError("setup_preconditioner reported an error")
            for _ in range(iterrfne):
                self.update()
        return error


class TetHpIns(_TetHp):
    NV = 4

    def __init__(self, mesh: TetMesh, log2p: int = 1, rho: float = 1.0, mu: float = 0.0, adis: float = 1.0,
                 fadd: float = 1.0, alpha=NSTAGE_ALPHA, beta=NSTAGE_BETA, cfl=TET_CFL_DEFAULT):
        self.nstage = len(alpha)
        prm = _lib.ParamsC()
        prm.physics = _lib.HPG_PHYSICS_INS
        prm.log2p = int(log2p)
        prm.rho, prm.mu, prm.adis, prm.fadd = rho, mu, adis, fadd
        self._stages(prm, alpha, beta, cfl)
        self._create(mesh, prm, log2p)


class TetHpCns(_TetHp):
    """tet_hp_cns: unknowns (p, u, v, w, RT), every boundary face an inflow face with the
    boundary state bc = (u, v, w, RT); Weiss-Smith preconditioner."""
    NV = 5

    def __init__(self, mesh: TetMesh, log2p: int = 1, mu: float = 0.0, adis: float = 1.0, fadd: float = 1.0,
                 gamma: float = 1.4, prandtl: float = 0.713, R: float = 287.058, bc=(0.0, 0.0, 0.0, 1.0),
                 alpha=NSTAGE_ALPHA, beta=NSTAGE_BETA, cfl=TET_CFL_DEFAULT):
        if len(bc) != 4:
            raise ValueError("bc is the inflow state (u, v, w, RT)")
        self.nstage = len(alpha)
        prm = _lib.ParamsC()
        prm.physics = _lib.HPG_PHYSICS_CNS
        prm.log2p = int(log2p)
        prm.rho, prm.mu, prm.adis, prm.fadd = 1.0, mu, adis, fadd
        prm.gamma, prm.prandtl, prm.Rgas = gamma, prandtl, R
        prm.bc[0] = 0.0
        for k in range(4):
            prm.bc[k + 1] = bc[k]
        self._stages(prm, alpha, beta, cfl)
        self._create(mesh, prm, log2p)

    def get_cns_blocks(self):
        """vpreconditioner (npnt, 5, 5) and epreconditioner (nseg, 5, 5)"""
        m = self.mesh
        vpc, epc = np.zeros((m.npnt, 5, 5)), np.zeros((m.nseg, 5, 5))
        self._check(self.L.hpg_tet_get_cns_blocks(self.h, _lib.ptr(vpc), _lib.ptr(epc)), "get_cns_blocks")
        return vpc, epc
