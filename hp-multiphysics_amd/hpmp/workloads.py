"""Seeded synthetic states for the tri_hp ins workloads.

Mode orders (polynomial degree) of the modal basis: vertex modes 1, side mode m
(0-based) m + 2, interior mode of eta-group g (2..sm) and index k: g + 1 + k
(tri_basis/ptvalues.cpp:84-195).  Essential boundary values (u, v on the
``inflow`` edges) are held at the free stream, so the discrete boundary flux of
the initial state sums to zero and the steady problem is solvable.
"""
from __future__ import annotations

import numpy as np

from .mesh import TriMesh


def mode_orders(p: int):
    sm = p - 1
    side = np.arange(sm) + 2
    inter = []
    for g in range(2, sm + 1):
        for k in range(sm + 1 - g):
            inter.append(g + 1 + k)
    return side, np.array(inter, dtype=np.int64)


def freestream(mesh: TriMesh, p: int, u0: float = 1.0, v0: float = 0.0):
    sm = p - 1
    im = (p - 1) * (p - 2) // 2
    v = np.zeros((mesh.npnt, 3)); v[:, 0] = u0; v[:, 1] = v0
    return v, np.zeros((mesh.nseg, sm, 3)), np.zeros((mesh.ntri, im, 3))


def _hold_boundary(mesh: TriMesh, p, v, s, vfree, sfree):
    bv = mesh.boundary_vertex_mask()
    bs = mesh.boundary_segment_mask()
    v[bv, :2] = vfree[bv, :2]
    if s.size:
        s[bs, :, :2] = sfree[bs, :, :2]


def noisy_freestream(mesh: TriMesh, p: int, seed: int = 0, scale: float = 1.0):
    """u = 1 free stream plus N(0, (scale*2^-k)^2) on every coefficient of order k."""
    rng = np.random.default_rng(seed)
    v, s, i = freestream(mesh, p)
    vf, sf, _ = freestream(mesh, p)
    side_k, int_k = mode_orders(p)
    v = v + scale * 0.5 * rng.standard_normal(v.shape)
    if s.size:
        s = s + scale * (2.0 ** -side_k)[None, :, None] * rng.standard_normal(s.shape)
    if i.size:
        i = i + scale * (2.0 ** -int_k)[None, :, None] * rng.standard_normal(i.shape)
    _hold_boundary(mesh, p, v, s, vf, sf)
    return v, s, i


def fourier_state(mesh: TriMesh, p: int, seed: int = 0, nmodes: int = 8, amp: float = 1.0, noise: float = 1e-3):
    """Free stream + a smooth perturbation (sum of `nmodes` seeded Fourier modes of
    wavenumber <= 8, amplitude `amp`, times a bubble that vanishes on the boundary)
    at the vertices, plus N(0, noise^2) on the modes of order >= 3."""
    rng = np.random.default_rng(seed)
    v, s, i = freestream(mesh, p)
    vf, sf, _ = freestream(mesh, p)
    x, y = mesh.pnts[:, 0], mesh.pnts[:, 1]
    bub = 16.0 * x * (1 - x) * y * (1 - y)
    for comp in (0, 1):
        f = np.zeros(mesh.npnt)
        for _ in range(nmodes):
            kx, ky = rng.integers(1, 9, size=2)
            ph = rng.uniform(0, 2 * np.pi, size=2)
            f += rng.uniform(-1, 1) * np.sin(np.pi * kx * x + ph[0]) * np.cos(np.pi * ky * y + ph[1])
        f *= amp / max(1e-30, np.abs(f).max())
        v[:, comp] += bub * f
    side_k, int_k = mode_orders(p)
    if s.size:
        s = s + noise * (side_k >= 3)[None, :, None] * rng.standard_normal(s.shape)
    if i.size:
        i = i + noise * (int_k >= 3)[None, :, None] * rng.standard_normal(i.shape)
    _hold_boundary(mesh, p, v, s, vf, sf)
    return v, s, i


# ---- tri_hp_cns states (primitive variables p, u, v, RT) ---------------------------

def cns_freestream(mesh: TriMesh, p: int, mach: float = 0.5, gamma: float = 1.4, pr: float = 1.0, rt: float = 1.0):
    """Uniform (p, u, v, RT) = (pr, mach * sqrt(gamma rt), 0, rt); returns (state, bc)."""
    sm = p - 1
    im = (p - 1) * (p - 2) // 2
    bc = (pr, mach * np.sqrt(gamma * rt), 0.0, rt)
    v = np.tile(np.array(bc), (mesh.npnt, 1))
    return (v, np.zeros((mesh.nseg, sm, 4)), np.zeros((mesh.ntri, im, 4))), bc


def _hold_cns_boundary(mesh, v, s, bc):
    bv = mesh.boundary_vertex_mask()
    bs = mesh.boundary_segment_mask()
    v[bv, 1:] = np.array(bc)[1:]
    if s.size:
        s[bs, :, 1:] = 0.0


def cns_gaussian_state(mesh: TriMesh, p: int, seed: int = 0, mach: float = 0.5, amp: float = 1e-2,
                       width: float = 0.05, nbumps: int = 8, noise: float = 0.0, gamma: float = 1.4):
    """Mach `mach` free stream plus seeded Gaussian density and temperature bumps
    (amplitude `amp`, width `width`, centres uniform in [0.15, 0.85]^2) at the
    vertices; p = rho RT.  Optional N(0, noise^2 2^-2k) on modes of order k >= 2.
    u, v, RT are held at the free stream on the boundary.  Returns (state, bc)."""
    rng = np.random.default_rng(seed)
    (v, s, i), bc = cns_freestream(mesh, p, mach, gamma)
    x, y = mesh.pnts[:, 0], mesh.pnts[:, 1]

    def bumps():
        f = np.zeros(mesh.npnt)
        for _ in range(nbumps):
            cx, cy = rng.uniform(0.15, 0.85, size=2)
            f += np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / width ** 2)
        return f

    rho = 1.0 + amp * bumps()
    rt = 1.0 + amp * bumps()
    v[:, 3] = rt
    v[:, 0] = rho * rt
    if noise:
        side_k, int_k = mode_orders(p)
        if s.size:
            s = s + noise * (2.0 ** -side_k)[None, :, None] * rng.standard_normal(s.shape)
        if i.size:
            i = i + noise * (2.0 ** -int_k)[None, :, None] * rng.standard_normal(i.shape)
    _hold_cns_boundary(mesh, v, s, bc)
    return (v, s, i), bc


def tet_solenoidal_state(mesh, p: int, seed: int = 0, nmodes: int = 16, kmax: int = 4):
    """Velocity u = curl psi, psi = sum of `nmodes` seeded Fourier modes A sin(k.x + phi)
    (so div u = 0 exactly), scaled to max |u| = 1 at the vertices, pressure 0.  Vertex
    values are point values; the p = 2 edge coefficient interpolates the edge midpoint:
    u(mid) = (u_a + u_b)/2 + c/4 for the edge mode lambda_a lambda_b.  Returns (v, e)
    in the layouts v(npnt, 4), e(nseg, p-1, 4)."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(-kmax, kmax + 1, size=(nmodes, 3)).astype(np.float64) * np.pi
    ks[np.all(ks == 0, axis=1)] = np.pi
    amps = rng.standard_normal((nmodes, 3))
    phis = rng.uniform(0, 2 * np.pi, nmodes)

    def vel(x):
        u = np.zeros_like(x)
        for k, a, ph in zip(ks, amps, phis):
            u += np.cross(k, a)[None, :] * np.cos(x @ k + ph)[:, None]
        return u

    uv = vel(mesh.pnts)
    scale = 1.0 / np.abs(uv).max()
    v = np.zeros((mesh.npnt, 4))
    v[:, :3] = uv * scale
    em = p - 1
    e = np.zeros((mesh.nseg, em, 4))
    if em:
        a, b = mesh.seg_pnt[:, 0], mesh.seg_pnt[:, 1]
        mid = 0.5 * (mesh.pnts[a] + mesh.pnts[b])
        e[:, 0, :3] = 4.0 * (vel(mid) * scale - 0.5 * (v[a, :3] + v[b, :3]))
    return v, e
