"""ctypes front end of the tet_hp_ins and tet_hp_cns part of the CPU oracle (test
infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  It drives the same liboracle.so as oracle.py; the C++ restatement
is oracle/otet_basis.cpp (tet_basis), oracle/otet_hp.cpp (tet_hp, tet_hp_ins) and
oracle/otet_cns.cpp (tet_hp_cns).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import lib as _base_lib

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_done = False

TET_CFL_DEFAULT = (2.2, 1.5, 1.0, 0.5)
NSTAGE_ALPHA = (0.25, 1.0 / 6.0, 0.375, 0.5, 1.0)
NSTAGE_BETA = (1.0, 0.0, 5.0 / 9.0, 0.0, 4.0 / 9.0)


def lib():
    global _done
    L = _base_lib()
    if not _done:
        L.otetb_info.argtypes = [C.c_int, _ip]
        L.otetb_get.argtypes = [C.c_int, C.c_char_p, C.c_void_p]
        L.otetb_get.restype = C.c_int
        L.otetb_proj.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.otetb_proj_val.argtypes = [C.c_int, _dp, _dp]
        L.otetb_proj_lin.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp]
        L.otetb_intgrt.argtypes = [C.c_int, _dp, _dp]
        L.otetb_intgrtrst.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        for nm in ("otetb_derivr", "otetb_derivs", "otetb_derivt"):
            getattr(L, nm).argtypes = [C.c_int, _dp, _dp]
        L.otetb_proj2d.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.otetb_intgrt2d.argtypes = [C.c_int, _dp, _dp]
        L.otet_create.restype = C.c_void_p
        L.otet_create.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _ip, _ip, C.c_int, _ip, _ip, _ip,
                                  _dp, C.c_int, C.c_int, _dp, _dp, _dp]
        L.otet_destroy.argtypes = [C.c_void_p]
        for nm in ("otet_set_state", "otet_get_state", "otet_get_res", "otet_set_res"):
            getattr(L, nm).argtypes = [C.c_void_p, _dp, _dp]
        L.otet_get_prcn.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.otet_set_tau.argtypes = [C.c_void_p, _dp]
        L.otet_rsdl.argtypes = [C.c_void_p, C.c_int]
        L.otet_setup_preconditioner.argtypes = [C.c_void_p]
        L.otet_setup_preconditioner.restype = C.c_int
        for nm in ("otet_minvrt", "otet_update", "otet_mg_restrict", "otet_mg_prolongate"):
            getattr(L, nm).argtypes = [C.c_void_p]
        L.otet_maxres.argtypes = [C.c_void_p, _dp]
        L.otet_maxres.restype = C.c_double
        L.otet_element_jacobian.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp]
        L.otet_log2p.argtypes = [C.c_void_p]
        L.otet_element_rsdl.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        L.otet_face_rsdl.argtypes = [C.c_void_p, C.c_int, _dp]
        L.otetc_create.restype = C.c_void_p
        L.otetc_create.argtypes = L.otet_create.argtypes
        L.otetc_destroy.argtypes = [C.c_void_p]
        for nm in ("otetc_set_state", "otetc_get_state", "otetc_get_res", "otetc_set_res"):
            getattr(L, nm).argtypes = [C.c_void_p, _dp, _dp]
        L.otetc_get_prcn.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.otetc_set_tau.argtypes = [C.c_void_p, _dp]
        L.otetc_rsdl.argtypes = [C.c_void_p, C.c_int]
        L.otetc_setup_preconditioner.argtypes = [C.c_void_p]
        L.otetc_setup_preconditioner.restype = C.c_int
        for nm in ("otetc_minvrt", "otetc_update", "otetc_mg_restrict", "otetc_mg_prolongate"):
            getattr(L, nm).argtypes = [C.c_void_p]
        L.otetc_maxres.argtypes = [C.c_void_p, _dp]
        L.otetc_maxres.restype = C.c_double
        L.otetc_element_jacobian.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp]
        L.otetc_log2p.argtypes = [C.c_void_p]
        L.otetc_element_rsdl.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        L.otetc_face_rsdl.argtypes = [C.c_void_p, C.c_int, _dp]
        L.otetc_tau_timestep.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, _dp, _dp, _dp]
        L.otetc_switch_variables.argtypes = [C.c_void_p, _dp, _dp]
        L.otetc_abs_jacobians.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp, _dp]
        L.otetc_flux_jacobians.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        _done = True
    return L


class TetBasis:
    """Tables and transforms of tet_basis for p in {1, 2}; Gauss arrays (gpx, gpy, gpz)."""

    def __init__(self, p: int):
        self.p = p
        info = np.zeros(6, dtype=np.int32)
        lib().otetb_info(p, info)
        self.p, self.em, self.tm, self.gp, self.G, self.tm2d = (int(x) for x in info)

    def table(self, name: str) -> np.ndarray:
        buf = np.zeros(4096)
        n = lib().otetb_get(self.p, name.encode(), buf.ctypes.data)
        if n < 0:
            raise KeyError(name)
        return buf[:n].copy()

    def proj(self, lin):
        f, dx, dy, dz = (np.zeros(self.G) for _ in range(4))
        lib().otetb_proj(self.p, np.ascontiguousarray(lin, dtype=np.float64), f, dx, dy, dz)
        return f, dx, dy, dz

    def proj_val(self, lin):
        f = np.zeros(self.G)
        lib().otetb_proj_val(self.p, np.ascontiguousarray(lin, dtype=np.float64), f)
        return f

    def proj_lin(self, u0, u1, u2, u3):
        f = np.zeros(self.G)
        lib().otetb_proj_lin(self.p, u0, u1, u2, u3, f)
        return f

    def intgrt(self, f):
        r = np.zeros(self.tm)
        lib().otetb_intgrt(self.p, r, np.ascontiguousarray(f, dtype=np.float64))
        return r

    def intgrtrst(self, dx, dy, dz, init=None):
        r = np.zeros(self.tm) if init is None else np.array(init, dtype=np.float64)
        lib().otetb_intgrtrst(self.p, r, *(np.ascontiguousarray(a, dtype=np.float64) for a in (dx, dy, dz)))
        return r

    def deriv(self, which: str, f, init=None):
        d = np.zeros(self.G) if init is None else np.array(init, dtype=np.float64)
        getattr(lib(), "otetb_deriv" + which)(self.p, np.ascontiguousarray(f, dtype=np.float64), d)
        return d

    def proj2d(self, lin):
        n = self.gp * self.gp
        f, dx, dy = np.zeros(n), np.zeros(n), np.zeros(n)
        lib().otetb_proj2d(self.p, np.ascontiguousarray(lin, dtype=np.float64), f, dx, dy)
        return f, dx, dy

    def intgrt2d(self, f):
        r = np.zeros(self.tm2d)
        lib().otetb_intgrt2d(self.p, r, np.ascontiguousarray(f, dtype=np.float64))
        return r


class TetHpIns:
    """tet_hp_ins on a straight tetrahedral mesh (hpmp.tetmesh.TetMesh)."""
    NV = 4

    def __init__(self, mesh, log2p=1, rho=1.0, mu=0.0, adis=1.0, fadd=1.0, alpha=NSTAGE_ALPHA,
                 beta=NSTAGE_BETA, cfl=TET_CFL_DEFAULT):
        L = lib()
        self.mesh = mesh
        self.log2pmax = int(log2p)
        self.em0 = (1 << self.log2pmax) - 1
        m = mesh
        self._keep = [np.ascontiguousarray(a) for a in (
            m.pnts.astype(np.float64), m.seg_pnt.astype(np.int32), m.tet_pnt.astype(np.int32),
            m.tet_seg.astype(np.int32), m.tet_sgn.astype(np.int32), m.bf_pnt.astype(np.int32),
            m.bf_seg.astype(np.int32), m.bf_sgn.astype(np.int32))]
        k = self._keep
        prm = np.array([rho, mu, adis, fadd], dtype=np.float64)
        cf = np.array(list(cfl) + [0.5] * 4, dtype=np.float64)[:4]
        self.h = L.otet_create(m.npnt, k[0].ravel(), m.nseg, k[1].ravel(), m.ntet, k[2].ravel(), k[3].ravel(),
                               k[4].ravel(), m.nbf, k[5].ravel(), k[6].ravel(), k[7].ravel(), prm, self.log2pmax,
                               len(alpha), np.array(alpha, dtype=np.float64), np.array(beta, dtype=np.float64), cf)
        if not self.h:
            raise RuntimeError("otet_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().otet_destroy(self.h)
            self.h = None

    def shapes(self):
        return (self.mesh.npnt, self.NV), (self.mesh.nseg, self.em0, self.NV)

    def zeros(self):
        return tuple(np.zeros(s) for s in self.shapes())

    def set_state(self, v, e):
        lib().otet_set_state(self.h, *(np.ascontiguousarray(a, dtype=np.float64) for a in (v, e)))

    def get_state(self):
        v, e = self.zeros()
        lib().otet_get_state(self.h, v, e)
        return v, e

    def get_res(self):
        v, e = self.zeros()
        lib().otet_get_res(self.h, v, e)
        return v, e

    def set_res(self, v, e):
        lib().otet_set_res(self.h, *(np.ascontiguousarray(a, dtype=np.float64) for a in (v, e)))

    def get_prcn(self):
        m = self.mesh
        vp, ep, ip, tau = (np.zeros((n, 4)) for n in (m.npnt, m.nseg, m.ntet, m.ntet))
        lib().otet_get_prcn(self.h, vp, ep, ip, tau)
        return vp, ep, ip, tau

    def set_tau(self, tau):
        lib().otet_set_tau(self.h, np.ascontiguousarray(tau, dtype=np.float64))

    def rsdl(self, stage=None):
        lib().otet_rsdl(self.h, len(NSTAGE_ALPHA) if stage is None else stage)

    def setup_preconditioner(self) -> int:
        return lib().otet_setup_preconditioner(self.h)

    def minvrt(self):
        lib().otet_minvrt(self.h)

    def update(self):
        lib().otet_update(self.h)

    def maxres(self):
        pv = np.zeros(4)
        return lib().otet_maxres(self.h, pv), pv

    def element_jacobian(self, tind, eps_r=1e-6, eps_a=1e-10):
        """tet_hp::element_jacobian (tet_hp/update.cpp:134-178), current level."""
        tm = 4 if self.log2p == 0 else 10
        K = np.zeros((tm * 4, tm * 4))
        lib().otet_element_jacobian(self.h, int(tind), eps_r, eps_a, K)
        return K

    def mg_restrict(self):
        lib().otet_mg_restrict(self.h)

    def mg_prolongate(self):
        lib().otet_mg_prolongate(self.h)

    @property
    def log2p(self):
        return lib().otet_log2p(self.h)

    def element_rsdl(self, tind, stage, uht):
        tm = uht.shape[1]
        lre, lim = np.zeros((4, tm)), np.zeros((4, tm))
        lib().otet_element_rsdl(self.h, tind, stage, np.ascontiguousarray(uht, dtype=np.float64), lre, lim)
        return lre, lim

    def face_rsdl(self, f):
        tf = 3 * (1 + ((1 << self.log2p) - 1))
        lf = np.zeros((4, tf))
        lib().otet_face_rsdl(self.h, f, lf)
        return lf


class TetHpCns:
    """tet_hp_cns (p, u, v, w, RT) on a straight tetrahedral mesh, every boundary face an
    inflow face with the boundary values bc = (unused, u, v, w, RT)."""
    NV = 5

    def __init__(self, mesh, log2p=1, mu=0.0, adis=1.0, fadd=1.0, gamma=1.4, prandtl=0.713, R=287.058,
                 bc=(0.0, 0.0, 0.0, 0.0, 1.0), preconditioner=1, alpha=NSTAGE_ALPHA, beta=NSTAGE_BETA,
                 cfl=TET_CFL_DEFAULT):
        L = lib()
        self.mesh = mesh
        self.log2pmax = int(log2p)
        self.em0 = (1 << self.log2pmax) - 1
        m = mesh
        self._keep = [np.ascontiguousarray(a) for a in (
            m.pnts.astype(np.float64), m.seg_pnt.astype(np.int32), m.tet_pnt.astype(np.int32),
            m.tet_seg.astype(np.int32), m.tet_sgn.astype(np.int32), m.bf_pnt.astype(np.int32),
            m.bf_seg.astype(np.int32), m.bf_sgn.astype(np.int32))]
        k = self._keep
        prm = np.array([mu, adis, fadd, gamma, prandtl, R, *bc, preconditioner], dtype=np.float64)
        cf = np.array(list(cfl) + [0.5] * 4, dtype=np.float64)[:4]
        self.h = L.otetc_create(m.npnt, k[0].ravel(), m.nseg, k[1].ravel(), m.ntet, k[2].ravel(), k[3].ravel(),
                                k[4].ravel(), m.nbf, k[5].ravel(), k[6].ravel(), k[7].ravel(), prm, self.log2pmax,
                                len(alpha), np.array(alpha, dtype=np.float64), np.array(beta, dtype=np.float64), cf)
        if not self.h:
            raise RuntimeError("otetc_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().otetc_destroy(self.h)
            self.h = None

    def shapes(self):
        return (self.mesh.npnt, self.NV), (self.mesh.nseg, self.em0, self.NV)

    def zeros(self):
        return tuple(np.zeros(s) for s in self.shapes())

    def set_state(self, v, e):
        lib().otetc_set_state(self.h, *(np.ascontiguousarray(a, dtype=np.float64) for a in (v, e)))

    def get_state(self):
        v, e = self.zeros()
        lib().otetc_get_state(self.h, v, e)
        return v, e

    def get_res(self):
        v, e = self.zeros()
        lib().otetc_get_res(self.h, v, e)
        return v, e

    def set_res(self, v, e):
        lib().otetc_set_res(self.h, *(np.ascontiguousarray(a, dtype=np.float64) for a in (v, e)))

    def get_prcn(self):
        """vprcn (npnt, 5), eprcn (nseg, 5), tau (ntet, 5, 5), vpc (npnt, 5, 5),
        epc (nseg, 5, 5), tstep (ntet,)"""
        m = self.mesh
        vp, ep = np.zeros((m.npnt, 5)), np.zeros((m.nseg, 5))
        tau, vpc, epc = np.zeros((m.ntet, 5, 5)), np.zeros((m.npnt, 5, 5)), np.zeros((m.nseg, 5, 5))
        ts = np.zeros(m.ntet)
        lib().otetc_get_prcn(self.h, vp, ep, tau, vpc, epc, ts)
        return vp, ep, tau, vpc, epc, ts

    def set_tau(self, tau):
        lib().otetc_set_tau(self.h, np.ascontiguousarray(tau, dtype=np.float64))

    def rsdl(self, stage=None):
        lib().otetc_rsdl(self.h, len(NSTAGE_ALPHA) if stage is None else stage)

    def setup_preconditioner(self) -> int:
        return lib().otetc_setup_preconditioner(self.h)

    def minvrt(self):
        lib().otetc_minvrt(self.h)

    def update(self):
        lib().otetc_update(self.h)

    def maxres(self):
        pv = np.zeros(5)
        return lib().otetc_maxres(self.h, pv), pv

    def element_jacobian(self, tind, eps_r=1e-6, eps_a=1e-10):
        """tet_hp::element_jacobian (tet_hp/update.cpp:134-178), current level."""
        tm = 4 if self.log2p == 0 else 10
        K = np.zeros((tm * 5, tm * 5))
        lib().otetc_element_jacobian(self.h, int(tind), eps_r, eps_a, K)
        return K

    def mg_restrict(self):
        lib().otetc_mg_restrict(self.h)

    def mg_prolongate(self):
        lib().otetc_mg_prolongate(self.h)

    @property
    def log2p(self):
        return lib().otetc_log2p(self.h)

    def element_rsdl(self, tind, stage, uht):
        tm = uht.shape[1]
        lre, lim = np.zeros((5, tm)), np.zeros((5, tm))
        lib().otetc_element_rsdl(self.h, tind, stage, np.ascontiguousarray(uht, dtype=np.float64), lre, lim)
        return lre, lim

    def face_rsdl(self, f):
        tf = 3 * (1 + ((1 << self.log2p) - 1))
        lf = np.zeros((5, tf))
        lib().otetc_face_rsdl(self.h, f, lf)
        return lf

    def tau_timestep(self, pvu, hmin, hmax):
        pinv, tau, ts = np.zeros(25), np.zeros(25), np.zeros(1)
        lib().otetc_tau_timestep(self.h, np.ascontiguousarray(pvu, dtype=np.float64), hmin, hmax, pinv, tau, ts)
        return pinv.reshape(5, 5), tau.reshape(5, 5), float(ts[0])

    def switch_variables(self, pvu, a):
        x = np.array(a, dtype=np.float64)
        lib().otetc_switch_variables(self.h, np.ascontiguousarray(pvu, dtype=np.float64), x)
        return x

    def abs_jacobians(self, pvu, b2):
        A, B, Cm = np.zeros(25), np.zeros(25), np.zeros(25)
        lib().otetc_abs_jacobians(self.h, np.ascontiguousarray(pvu, dtype=np.float64), float(b2), A, B, Cm)
        return A.reshape(5, 5), B.reshape(5, 5), Cm.reshape(5, 5)

    def flux_jacobians(self, w):
        A, B, Cm = np.zeros(25), np.zeros(25), np.zeros(25)
        lib().otetc_flux_jacobians(self.h, np.ascontiguousarray(w, dtype=np.float64), A, B, Cm)
        return A.reshape(5, 5), B.reshape(5, 5), Cm.reshape(5, 5)
