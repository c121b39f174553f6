"""Host-side mirror of tet_hp / tet_hp_ins (tet_hp/tet_hp.h:229-273,
tet_hp/ins/tet_hp_ins.h) over the hpg_tet_* C-ABI of libhpgpu.so.

Same method names and argument meaning as the reference's virtual interface:
rsdl(stage), setup_preconditioner() -> error count, minvrt(), update(), maxres(),
mg_restrict(), mg_prolongate().  State is exchanged in the reference layouts
v(npnt, 4) and e(nseg, em0, 4) (tet_hp/tet_hp.h:89-94).  There is no CPU path:
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


class TetHpIns:
    NV = 4

    def __init__(self, mesh: TetMesh, log2p: int = 1, rho: float = 1.0, mu: float = 0.0, adis: float = 1.0,
                 fadd: float = 1.0, alpha=NSTAGE_ALPHA, beta=NSTAGE_BETA, cfl=TET_CFL_DEFAULT):
        self.L = _lib.load()
        self.mesh = mesh
        self.log2pmax = int(log2p)
        self.em0 = (1 << self.log2pmax) - 1
        self.nstage = len(alpha)
        m = mesh
        self._keep = [np.ascontiguousarray(a) for a in (
            m.pnts.astype(np.float64), m.seg_pnt.astype(np.int32), m.tet_pnt.astype(np.int32),
            m.tet_seg.astype(np.int32), m.tet_sgn.astype(np.int32), m.bf_pnt.astype(np.int32),
            m.bf_seg.astype(np.int32), m.bf_sgn.astype(np.int32))]
        cm = _lib.TetMeshC(m.npnt, m.nseg, m.ntet, m.nbf, *(C.c_void_p(a.ctypes.data) for a in self._keep))
        prm = _lib.ParamsC()
        prm.physics = _lib.HPG_PHYSICS_INS
        prm.log2p = self.log2pmax
        prm.rho, prm.mu, prm.adis, prm.fadd = rho, mu, adis, fadd
        prm.nstage = self.nstage
        for s in range(self.nstage):
            prm.alpha[s] = alpha[s]
            prm.beta[s] = beta[s]
        for lvl in range(4):
            prm.cfl[lvl] = cfl[lvl] if lvl < len(cfl) else 0.5
        h = C.c_void_p()
        rc = self.L.hpg_tet_create(C.byref(cm), C.byref(prm), C.byref(h))
        if rc != 0 or not h.value:
            raise _lib.HpgError(f"hpg_tet_create failed ({rc})")
        self.h = h

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
        m = self.mesh
        out = tuple(np.zeros((n, 4)) for n in (m.npnt, m.nseg, m.ntet, m.ntet))
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
