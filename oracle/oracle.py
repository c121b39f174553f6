"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module.  It drives oracle/_build/liboracle.so, the C++ restatement of the
reference's tri_basis and tri_hp/tri_hp_ins residual/relaxation path (see the
headers of oracle/otri_basis.h and oracle/otri_hp_ins.h for the file:line map).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def set_threads(n: int) -> int:
    """OpenMP threads of the oracle's element loops (results do not depend on it)."""
    return int(lib().oracle_set_threads(int(n)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_recur.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.oracle_gauss.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.obasis_info.argtypes = [C.c_int, _ip]
        L.obasis_get.argtypes = [C.c_int, C.c_char_p, C.c_void_p]
        L.obasis_get.restype = C.c_int
        L.obasis_ptvalues_deriv.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]
        for nm in ("obasis_proj",):
            getattr(L, nm).argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.obasis_proj_val.argtypes = [C.c_int, _dp, _dp]
        L.obasis_proj_lin.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.obasis_derivr.argtypes = [C.c_int, _dp, _dp]
        L.obasis_derivs.argtypes = [C.c_int, _dp, _dp]
        L.obasis_intgrt.argtypes = [C.c_int, _dp, _dp]
        L.obasis_intgrtrs.argtypes = [C.c_int, _dp, _dp, _dp]
        L.obasis_proj1d.argtypes = [C.c_int, _dp, _dp, _dp]
        L.obasis_intgrt1d.argtypes = [C.c_int, _dp, _dp]
        L.oracle_dpbtrsnu2.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int]
        L.otri_create.restype = C.c_void_p
        L.otri_create.argtypes = [C.c_int, _dp, C.c_int, _ip, C.c_int, _ip, _ip, _ip, C.c_int, _ip, _ip, _ip,
                                  _dp, C.c_int, C.c_int, _dp, _dp, _dp]
        L.otri_destroy.argtypes = [C.c_void_p]
        for nm in ("otri_set_state", "otri_get_state", "otri_get_res", "otri_set_res", "otri_get_prcn"):
            getattr(L, nm).argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.otri_set_tprcn.argtypes = [C.c_void_p, _dp]
        L.otri_assemble_prcn.argtypes = [C.c_void_p]
        L.otri_rsdl.argtypes = [C.c_void_p, C.c_int]
        L.otri_setup_preconditioner.argtypes = [C.c_void_p]
        L.otri_setup_preconditioner.restype = C.c_int
        for nm in ("otri_minvrt", "otri_update", "otri_mg_restrict", "otri_mg_prolongate"):
            getattr(L, nm).argtypes = [C.c_void_p]
        L.otri_maxres.argtypes = [C.c_void_p, _dp]
        L.otri_maxres.restype = C.c_double
        L.otri_log2p.argtypes = [C.c_void_p]
        L.otri_cycle.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.otri_cycle.restype = C.c_double
        L.otri_element_rsdl.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        L.otri_element_jacobian.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp]
        L.otri_link_fine.argtypes = [C.c_void_p, C.c_void_p, _ip, _dp, _ip, _dp]
        L.otri_set_curved.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
        L.otri_set_edge_flux.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
        L.otri_set_periodic.argtypes = [C.c_void_p, C.c_int, _ip]
        L.otri_edge_flux.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.ocns_link_fine.argtypes = [C.c_void_p, C.c_void_p, _ip, _dp, _ip, _dp]
        L.otri_get_mg.argtypes = [C.c_void_p, _dp, _dp]
        L.ocns_get_mg.argtypes = [C.c_void_p, _dp, _dp]
        L.ocns_create.restype = C.c_void_p
        L.ocns_create.argtypes = L.otri_create.argtypes
        L.ocns_destroy.argtypes = [C.c_void_p]
        for nm in ("ocns_set_state", "ocns_get_state", "ocns_get_res", "ocns_set_res"):
            getattr(L, nm).argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.ocns_get_prcn.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.ocns_rsdl.argtypes = [C.c_void_p, C.c_int]
        L.ocns_set_vpc.argtypes = [C.c_void_p, _dp]
        L.ocns_setup_preconditioner.argtypes = [C.c_void_p]
        L.ocns_setup_preconditioner.restype = C.c_int
        for nm in ("ocns_minvrt", "ocns_update", "ocns_mg_restrict", "ocns_mg_prolongate",
                   "ocns_modify_boundary_residual", "ocns_project_new_variables"):
            getattr(L, nm).argtypes = [C.c_void_p]
        L.ocns_maxres.argtypes = [C.c_void_p, _dp]
        L.ocns_maxres.restype = C.c_double
        L.ocns_log2p.argtypes = [C.c_void_p]
        L.ocns_cycle.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ocns_cycle.restype = C.c_double
        L.ocns_element_rsdl.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp]
        L.ocns_element_jacobian.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, _dp]
        L.ocns_calculate_tau.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp, _dp]
        L.ocns_spectral_radius.argtypes = [_dp]
        L.ocns_spectral_radius.restype = C.c_double
        L.oracle_set_threads(1)  # one thread unless a caller asks for more
        _lib = L
    return _lib


class Basis:
    """Tables and transforms of the oracle's tri_basis<p,0>."""

    def __init__(self, p: int):
        self.p = p
        info = np.zeros(7, dtype=np.int32)
        lib().obasis_info(p, info)
        self.sm, self.im, self.tm, self.bm, self.nmodx, self.gpx, self.gpn = (int(v) for v in info)

    def table(self, name: str) -> np.ndarray:
        n = lib().obasis_get(self.p, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().obasis_get(self.p, name.encode(), out.ctypes.data_as(C.c_void_p))
        return out

    def ptvalues_deriv(self, x, eta):
        a = np.zeros(self.nmodx); b = np.zeros(self.nmodx); c = np.zeros(self.tm); d = np.zeros(self.tm)
        lib().obasis_ptvalues_deriv(self.p, x, eta, a, b, c, d)
        return a, b, c, d

    def proj(self, lin):
        f = np.zeros((self.gpx, self.gpn)); dx = np.zeros_like(f); dy = np.zeros_like(f)
        lib().obasis_proj(self.p, np.ascontiguousarray(lin, dtype=np.float64), f, dx, dy)
        return f, dx, dy

    def proj_val(self, lin):
        f = np.zeros((self.gpx, self.gpn))
        lib().obasis_proj_val(self.p, np.ascontiguousarray(lin, dtype=np.float64), f)
        return f

    def proj_lin(self, u1, u2, u3):
        f = np.zeros((self.gpx, self.gpn))
        lib().obasis_proj_lin(self.p, u1, u2, u3, f)
        return f

    def derivr(self, f):
        out = np.zeros((self.gpx, self.gpn))
        lib().obasis_derivr(self.p, np.ascontiguousarray(f, dtype=np.float64), out)
        return out

    def derivs(self, f):
        out = np.zeros((self.gpx, self.gpn))
        lib().obasis_derivs(self.p, np.ascontiguousarray(f, dtype=np.float64), out)
        return out

    def intgrt(self, f):
        out = np.zeros(self.tm)
        lib().obasis_intgrt(self.p, out, np.ascontiguousarray(f, dtype=np.float64))
        return out

    def intgrtrs(self, dx, dy, init=None):
        out = np.zeros(self.tm) if init is None else np.array(init, dtype=np.float64)
        lib().obasis_intgrtrs(self.p, out, np.ascontiguousarray(dx, dtype=np.float64),
                              np.ascontiguousarray(dy, dtype=np.float64))
        return out

    def proj1d(self, lin):
        f = np.zeros(self.gpx); dx = np.zeros(self.gpx)
        lib().obasis_proj1d(self.p, np.ascontiguousarray(lin, dtype=np.float64), f, dx)
        return f, dx

    def intgrt1d(self, f):
        out = np.zeros(self.sm + 2)
        lib().obasis_intgrt1d(self.p, out, np.ascontiguousarray(f, dtype=np.float64))
        return out


def recur(n, ipoly, al=0.0, be=0.0):
    a = np.zeros(n); b = np.zeros(n)
    lib().oracle_recur(n, ipoly, al, be, a, b)
    return a, b


def gauss(n, a, b):
    x = np.zeros(n); w = np.zeros(n)
    lib().oracle_gauss(n, np.ascontiguousarray(a[:n]), np.ascontiguousarray(b[:n]), x, w)
    return x, w


NSTAGE_ALPHA = (0.25, 1.0 / 6.0, 0.375, 0.5, 1.0)
NSTAGE_BETA = (1.0, 0.0, 5.0 / 9.0, 0.0, 4.0 / 9.0)


class TriHpIns:
    """Oracle tri_hp_ins on a TriMesh (see hpmp.mesh); reference array layouts."""

    NV = 3

    def __init__(self, mesh, log2p=2, rho=1.0, mu=0.0, adis=1.0, bc_pressure=0.0, fadd=1.0,
                 alpha=NSTAGE_ALPHA, beta=NSTAGE_BETA, cfl=(2.5, 1.5, 1.0)):
        self.mesh = mesh
        self.log2pmax = log2p
        self.p = 1 << log2p
        self.sm0 = self.p - 1
        self.im0 = (self.p - 1) * (self.p - 2) // 2
        prm = np.array([rho, mu, adis, bc_pressure, fadd], dtype=np.float64)
        al = np.array(alpha, dtype=np.float64)
        be = np.array(beta, dtype=np.float64)
        cf = np.array(list(cfl) + [0.5] * 4, dtype=np.float64)[:4]
        m = mesh
        self._keep = [np.ascontiguousarray(x) for x in (
            m.pnts.astype(np.float64), m.seg_pnt.astype(np.int32), m.tri_pnt.astype(np.int32),
            m.tri_seg.astype(np.int32), m.tri_sgn.astype(np.int32), m.ebdry_start.astype(np.int32),
            m.ebdry_seg.astype(np.int32), m.ebdry_essential.astype(np.int32))]
        k = self._keep
        self.h = lib().otri_create(m.npnt, k[0], m.nseg, k[1], m.ntri, k[2], k[3], k[4], m.nebd, k[5], k[6],
                                   k[7], prm, log2p, len(al), al, be, cf)
        if not self.h:
            raise RuntimeError("otri_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().otri_destroy(self.h)
            self.h = None

    def shapes(self):
        m = self.mesh
        return ((m.npnt, self.NV), (m.nseg, self.sm0, self.NV), (m.ntri, self.im0, self.NV))

    def _alloc(self):
        return [np.zeros(s) for s in self.shapes()]

    @staticmethod
    def _ptr(a):
        return a if a.size else np.zeros(1)

    def set_state(self, v, s, i):
        lib().otri_set_state(self.h, np.ascontiguousarray(v, dtype=np.float64),
                             self._ptr(np.ascontiguousarray(s, dtype=np.float64)),
                             self._ptr(np.ascontiguousarray(i, dtype=np.float64)))

    def get_state(self):
        v, s, i = self._alloc()
        lib().otri_get_state(self.h, v, self._ptr(s), self._ptr(i))
        return v, s, i

    def get_res(self):
        v, s, i = self._alloc()
        lib().otri_get_res(self.h, v, self._ptr(s), self._ptr(i))
        return v, s, i

    def set_periodic(self, pairs):
        """Periodic pairs of edge boundaries (the reference's prdc pairing)."""
        pr = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        if lib().otri_set_periodic(self.h, pr.size // 2, pr) != 0:
            raise RuntimeError("otri_set_periodic failed")

    def set_edge_flux(self, types, ibc):
        """Flux type of each edge boundary (0 generic, 1 euler, 2 characteristic) and the
        freestream state (u, v, p)."""
        t = np.ascontiguousarray(types, dtype=np.int32)
        lib().otri_set_edge_flux(self.h, t.size, t, np.ascontiguousarray(ibc, dtype=np.float64))

    def set_curved(self, sides, crds):
        """Curved boundary sides: crds (nside, sm0, 2), the coordinate side modes of each
        listed side in its own direction (hp_edge_bdry::crds)."""
        sides = np.ascontiguousarray(sides, dtype=np.int32).reshape(-1)
        crds = np.ascontiguousarray(crds, dtype=np.float64).reshape(sides.size, self.sm0, 2)
        self._curved = (sides, crds)
        lib().otri_set_curved(self.h, sides.size, sides if sides.size else np.zeros(1, np.int32),
                              crds.reshape(-1) if crds.size else np.zeros(1))

    def set_res(self, v, s, i):
        lib().otri_set_res(self.h, np.ascontiguousarray(v, dtype=np.float64),
                           self._ptr(np.ascontiguousarray(s, dtype=np.float64)),
                           self._ptr(np.ascontiguousarray(i, dtype=np.float64)))

    def get_prcn(self):
        m = self.mesh
        vp = np.zeros((m.npnt, 3)); sp = np.zeros((m.nseg, 3)); tp = np.zeros((m.ntri, 3))
        lib().otri_get_prcn(self.h, vp, sp, tp)
        return vp, sp, tp

    def set_tprcn(self, tp):
        lib().otri_set_tprcn(self.h, np.ascontiguousarray(tp, dtype=np.float64))

    def assemble_prcn(self):
        lib().otri_assemble_prcn(self.h)

    def rsdl(self, stage=None):
        lib().otri_rsdl(self.h, len(NSTAGE_ALPHA) if stage is None else stage)

    def setup_preconditioner(self) -> int:
        return lib().otri_setup_preconditioner(self.h)

    def minvrt(self):
        lib().otri_minvrt(self.h)

    def update(self):
        lib().otri_update(self.h)

    def maxres(self):
        pv = np.zeros(3)
        e = lib().otri_maxres(self.h, pv)
        return e, pv

    def mg_restrict(self):
        lib().otri_mg_restrict(self.h)

    def mg_prolongate(self):
        lib().otri_mg_prolongate(self.h)

    @property
    def log2p(self):
        return lib().otri_log2p(self.h)

    def cycle(self, vw=1, mglvls=1, itercrsn=1, iterrfne=0):
        return lib().otri_cycle(self.h, vw, mglvls, itercrsn, iterrfne)

    def link_fine(self, fine, cc_tri, cc_wt, fc_tri, fc_wt):
        """Make this (coarse mesh) level the h-multigrid neighbour of `fine`."""
        self._fine = fine
        self._conn = [np.ascontiguousarray(cc_tri, dtype=np.int32), np.ascontiguousarray(cc_wt, dtype=np.float64),
                      np.ascontiguousarray(fc_tri, dtype=np.int32), np.ascontiguousarray(fc_wt, dtype=np.float64)]
        lib().otri_link_fine(self.h, fine.h, *self._conn)

    def get_mg(self):
        """(res0 at the vertices, vug_frst): the multigrid entry residual and state."""
        r0 = np.zeros((self.mesh.npnt, 3)); vug = np.zeros((self.mesh.npnt, 3))
        lib().otri_get_mg(self.h, r0, vug)
        return r0, vug

    def element_rsdl(self, tind, stage, uht):
        tm = (self.p + 1) * (self.p + 2) // 2 if self.log2p == self.log2pmax else None
        tm = uht.shape[1]
        lr = np.zeros((3, tm)); li = np.zeros((3, tm))
        lib().otri_element_rsdl(self.h, tind, stage, np.ascontiguousarray(uht, dtype=np.float64), lr, li)
        return lr, li

    def element_jacobian(self, tind, eps_r=1e-8, eps_a=1e-8):
        tm = (self.p + 1) * (self.p + 2) // 2
        K = np.zeros((tm * 3, tm * 3))
        lib().otri_element_jacobian(self.h, tind, eps_r, eps_a, K)
        return K


CNS_CFL_DEFAULT = (2.5, 1.5, 1.0)


class TriHpCns:
    """Oracle tri_hp_cns on a TriMesh; variables (p, u, v, RT); reference array layouts.
    Boundaries whose essential mask is 0b1110 are inflow boundaries with the constant
    values `bc` = (p, u, v, RT)."""

    NV = 4

    def __init__(self, mesh, log2p=2, mu=1e-3, adis=1.0, gamma=1.4, prandtl=0.713, Rgas=1.0, fadd=1.0,
                 bc=(1.0, 0.5, 0.0, 1.0), alpha=NSTAGE_ALPHA, beta=NSTAGE_BETA, cfl=CNS_CFL_DEFAULT):
        self.mesh = mesh
        self.log2pmax = log2p
        self.p = 1 << log2p
        self.sm0 = self.p - 1
        self.im0 = (self.p - 1) * (self.p - 2) // 2
        prm = np.array([mu, adis, gamma, prandtl, Rgas, fadd, *bc], dtype=np.float64)
        al = np.array(alpha, dtype=np.float64)
        be = np.array(beta, dtype=np.float64)
        cf = np.array(list(cfl) + [0.5] * 4, dtype=np.float64)[:4]
        m = mesh
        self._keep = [np.ascontiguousarray(x) for x in (
            m.pnts.astype(np.float64), m.seg_pnt.astype(np.int32), m.tri_pnt.astype(np.int32),
            m.tri_seg.astype(np.int32), m.tri_sgn.astype(np.int32), m.ebdry_start.astype(np.int32),
            m.ebdry_seg.astype(np.int32), m.ebdry_essential.astype(np.int32))]
        k = self._keep
        self.h = lib().ocns_create(m.npnt, k[0], m.nseg, k[1], m.ntri, k[2], k[3], k[4], m.nebd, k[5], k[6],
                                   k[7], prm, log2p, len(al), al, be, cf)
        if not self.h:
            raise RuntimeError("ocns_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ocns_destroy(self.h)
            self.h = None

    def shapes(self):
        m = self.mesh
        return ((m.npnt, self.NV), (m.nseg, self.sm0, self.NV), (m.ntri, self.im0, self.NV))

    def _alloc(self):
        return [np.zeros(s) for s in self.shapes()]

    @staticmethod
    def _ptr(a):
        return a if a.size else np.zeros(1)

    def set_state(self, v, s, i):
        lib().ocns_set_state(self.h, np.ascontiguousarray(v, dtype=np.float64),
                             self._ptr(np.ascontiguousarray(s, dtype=np.float64)),
                             self._ptr(np.ascontiguousarray(i, dtype=np.float64)))

    def get_state(self):
        v, s, i = self._alloc()
        lib().ocns_get_state(self.h, v, self._ptr(s), self._ptr(i))
        return v, s, i

    def get_res(self):
        v, s, i = self._alloc()
        lib().ocns_get_res(self.h, v, self._ptr(s), self._ptr(i))
        return v, s, i

    def set_res(self, v, s, i):
        lib().ocns_set_res(self.h, np.ascontiguousarray(v, dtype=np.float64),
                           self._ptr(np.ascontiguousarray(s, dtype=np.float64)),
                           self._ptr(np.ascontiguousarray(i, dtype=np.float64)))

    def get_prcn(self):
        m = self.mesh
        out = [np.zeros((m.npnt, 4)), np.zeros((m.nseg, 4)), np.zeros((m.ntri, 4)),
               np.zeros((m.npnt, 4, 4)), np.zeros((m.nseg, 4, 4)), np.zeros((m.ntri, 4, 4))]
        lib().ocns_get_prcn(self.h, *out)
        return out

    def set_vpc(self, vpc):
        """Overwrite the vertex 4x4 preconditioner blocks (npnt x 4 x 4, row-major)."""
        lib().ocns_set_vpc(self.h, np.ascontiguousarray(vpc, dtype=np.float64).reshape(-1))

    def rsdl(self, stage=None):
        lib().ocns_rsdl(self.h, len(NSTAGE_ALPHA) if stage is None else stage)

    def setup_preconditioner(self) -> int:
        return lib().ocns_setup_preconditioner(self.h)

    def minvrt(self):
        lib().ocns_minvrt(self.h)

    def modify_boundary_residual(self):
        lib().ocns_modify_boundary_residual(self.h)

    def project_new_variables(self):
        lib().ocns_project_new_variables(self.h)

    def update(self):
        lib().ocns_update(self.h)

    def maxres(self):
        pv = np.zeros(4)
        e = lib().ocns_maxres(self.h, pv)
        return e, pv

    def mg_restrict(self):
        lib().ocns_mg_restrict(self.h)

    def mg_prolongate(self):
        lib().ocns_mg_prolongate(self.h)

    @property
    def log2p(self):
        return lib().ocns_log2p(self.h)

    def cycle(self, vw=1, mglvls=1, itercrsn=1, iterrfne=0):
        return lib().ocns_cycle(self.h, vw, mglvls, itercrsn, iterrfne)

    def element_rsdl(self, tind, stage, uht):
        tm = uht.shape[1]
        lr = np.zeros((4, tm)); li = np.zeros((4, tm))
        lib().ocns_element_rsdl(self.h, tind, stage, np.ascontiguousarray(uht, dtype=np.float64), lr, li)
        return lr, li

    def element_jacobian(self, tind, eps_r=1e-8, eps_a=1e-8):
        tm = (self.p + 1) * (self.p + 2) // 2
        K = np.zeros((tm * 4, tm * 4))
        lib().ocns_element_jacobian(self.h, tind, eps_r, eps_a, K)
        return K

    def link_fine(self, fine, cc_tri, cc_wt, fc_tri, fc_wt):
        """Make this (coarse mesh) level the h-multigrid neighbour of `fine`."""
        self._fine = fine
        self._conn = [np.ascontiguousarray(cc_tri, dtype=np.int32), np.ascontiguousarray(cc_wt, dtype=np.float64),
                      np.ascontiguousarray(fc_tri, dtype=np.int32), np.ascontiguousarray(fc_wt, dtype=np.float64)]
        lib().ocns_link_fine(self.h, fine.h, *self._conn)

    def get_mg(self):
        """(res0 at the vertices, vug_frst): the multigrid entry residual and state."""
        r0 = np.zeros((self.mesh.npnt, 4)); vug = np.zeros((self.mesh.npnt, 4))
        lib().ocns_get_mg(self.h, r0, vug)
        return r0, vug

    def calculate_tau(self, pvu, h):
        P = np.zeros(16); T = np.zeros(16); ts = np.zeros(1)
        lib().ocns_calculate_tau(self.h, np.ascontiguousarray(pvu, dtype=np.float64), h, P, T, ts)
        return P.reshape(4, 4), T.reshape(4, 4), float(ts[0])


def spectral_radius4(A):
    return lib().ocns_spectral_radius(np.ascontiguousarray(A, dtype=np.float64).ravel())


def edge_flux(ftype, u, nrm, side_length, p, rho, mu, ub):
    """One boundary point's flux of tri_hp_ins (generic 0, euler 1, characteristic 2)."""
    out = np.zeros(3)
    lib().otri_edge_flux(int(ftype), np.ascontiguousarray(u, dtype=np.float64), np.ascontiguousarray(nrm, dtype=np.float64),
                         float(side_length), int(p), float(rho), float(mu), np.ascontiguousarray(ub, dtype=np.float64), out)
    return out
