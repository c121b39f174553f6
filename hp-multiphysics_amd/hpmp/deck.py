"""Run a reference tri_hp_ins input deck (``.inpt``) on the GPU path.

The deck's keys are read as the reference reads them (``hpmp.input_map``: ``b0_``
prefixes, formula values) and mapped onto the library's setters:

* the mesh ``b0_mesh`` (a ``.grd`` file, relative to the deck's folder; a deck that names
  a mesh missing from its folder takes the ``mesh`` argument instead);
* ``b0_rho``, ``b0_mu``, ``log2p``, ``nstage``/``alpha``/``beta``, ``cfl``/``cfls``
  (tri_hp/allocate.cpp:240-290) through TriHpIns.from_input_map;
* every edge boundary's ``b0_s<id>_hp_type`` (tri_hp/ins/getnewbdry.cpp:70-110): inflow,
  outflow and symmetry as essential masks of the generic flux, euler and characteristic
  as flux types; ``b0_s<id>_type: prdc`` boundaries of one id pair periodically;
* a ``b0_s<id>_curved: 1`` boundary of ``b0_s<id>_type: symbolic`` with a level set
  ``b0_s<id>_h`` (and its gradient ``_dhdx0``, ``_dhdx1``) is curved at p >= 2: points go
  onto h = 0 by Newton steps along the gradient (edge_bdry::mvpttobdry of the symbolic
  boundary) and hpmp.curved.side_modes forms the side modes;
* ``ibc: freestream`` (``flowspeed``, ``flowangle`` in degrees, ``perturb_amplitude``;
  tri_hp/ins/getnewibc.cpp:16-51) gives the boundary state and, at time zero, the
  initial state u = speed cos a + amp x (1 - x), v = speed sin a, p = 0;
* the cycle: ``ncycle``, ``vwcycle``, ``itercrsn``, ``iterrfne``, ``extra_finest_levels``
  and ``ngrid`` (coarser meshes from hpmp.coarsen) as block::cycle
  (tri_mesh/blocks.cpp:1267-1348), stopping at ``relative_tolerance`` or
  ``absolute_tolerance`` of the max residual.

What the library has no counterpart for raises ValueError naming the key: other
physics, boundary data given by symbolic functions (applied_stress, a symbolic ibc),
mesh movement, periodic meshes with h-multigrid levels.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import grd, multigrid as MG
from .input_map import InputMap, evaluate
from .tri_hp import TriHpIns

INS_TYPES = ("inflow", "outflow", "symmetry", "euler", "characteristic")


def _resolve_mesh(inmap: InputMap, deck_dir: str, prefix: str, mesh: str | None) -> str:
    if mesh is not None:
        return mesh
    name = inmap.get(f"{prefix}_mesh", None)
    if name is None:
        raise ValueError(f"{prefix}_mesh is missing")
    name = os.path.expandvars(name.strip())
    for cand in (name, name + ".grd"):
        p = cand if os.path.isabs(cand) else os.path.join(deck_dir, cand)
        if os.path.isfile(p):
            return p
    raise FileNotFoundError(f"{prefix}_mesh '{name}' not found next to the deck; pass mesh=")


def _level_set_projector(inmap: InputMap, key: str):
    """to_boundary of a symbolic boundary: Newton steps x <- x - h grad h / |grad h|^2 on
    the level set h(x0, x1) = 0 with the deck's formulas."""
    h = inmap.get(f"{key}_h", None)
    g0, g1 = inmap.get(f"{key}_dhdx0", None), inmap.get(f"{key}_dhdx1", None)
    if h is None or g0 is None or g1 is None:
        raise ValueError(f"{key}: a curved symbolic boundary needs _h, _dhdx0 and _dhdx1")

    def ev(expr, x):
        return evaluate(expr, lambda nm: {"x0": x[0], "x1": x[1]}.get(nm, inmap.value(nm)))

    def to_boundary(pts):
        out = np.array(pts, dtype=np.float64)
        for p in out:
            for _ in range(50):
                hv = ev(h, p)
                gr = np.array([ev(g0, p), ev(g1, p)])
                step = hv * gr / (gr @ gr)
                p -= step
                if np.abs(step).max() < 1e-15 * (1.0 + np.abs(p).max()):
                    break
        return out
    return to_boundary


class Deck:
    """A tri_hp_ins deck: the mesh, boundary setup and cycle settings read from the deck
    (host only); ``build`` creates the GPU contexts and ``run`` performs the cycles."""

    def __init__(self, path: str, mesh: str | None = None, prefix: str = "b0"):
        self.path = path
        self.inmap = im = InputMap.from_file(path)
        self.prefix = prefix
        typ = (im.get(f"{prefix}_type", "") or "").split()
        if typ[:1] != ["ins"]:
            raise ValueError(f"{prefix}_type: '{' '.join(typ)}' (only ins decks run on this path)")
        if im.get("mesh_movement", None) is not None:
            raise ValueError("mesh_movement: moving meshes have no counterpart")
        if im.getf("dtinv", 0.0, prefix="") != 0.0:
            raise ValueError("dtinv: time-accurate runs have no counterpart (steady decks only)")
        deck_dir = os.path.dirname(os.path.abspath(path))
        self.mesh_path = _resolve_mesh(im, deck_dir, prefix, mesh)
        probe = grd.read_grd(self.mesh_path, 0)
        ids = list(probe.ebdry_ids)
        masks, types, prdc, curved = {}, {}, [], []
        for b, i in enumerate(ids):
            key = f"{prefix}_s{i}"
            btype = (im.get(f"{key}_type", "plain") or "plain").split()[0]
            if btype == "prdc":
                prdc.append(b)
                continue
            t = (im.get(f"{key}_hp_type", "plain") or "plain").split()[0]
            if t not in INS_TYPES:
                raise ValueError(f"{key}_hp_type: '{t}' is not supported (supported: {', '.join(INS_TYPES)})")
            if t == "inflow":
                masks[i] = 0b011
            elif t == "symmetry":
                d = im.geti(f"{key}_dir", 0, prefix="")
                if d not in (0, 1):
                    raise ValueError(f"{key}_dir must be 0 or 1")
                masks[i] = 1 << d
            else:
                masks[i] = 0
            types[b] = {"euler": 1, "characteristic": 2}.get(t, 0)
            if im.geti(f"{key}_curved", 0, prefix="") and btype == "symbolic":
                curved.append((b, key))
        self.mesh = m = grd.read_grd(self.mesh_path, masks)
        # periodic boundaries of one id pair in mesh order
        by_id = {}
        for b in prdc:
            by_id.setdefault(ids[b], []).append(b)
        pairs = []
        for i, bs in by_id.items():
            if len(bs) != 2:
                raise ValueError(f"{prefix}_s{i}_type: prdc needs exactly two edge boundaries of id {i}")
            pairs.append(tuple(bs))
        self.pairs = pairs
        # the freestream ibc
        ibc = (im.get("ibc", "freestream") or "freestream").split()[0]
        if ibc != "freestream":
            raise ValueError(f"ibc: '{ibc}' (only the freestream ibc has a counterpart)")
        speed = im.getf("flowspeed", 1.0, prefix="")
        angle = im.getf("flowangle", 0.0, prefix="") * math.pi / 180.0
        self.amp = im.getf("perturb_amplitude", 0.0, prefix="")
        self.ibc = (speed * math.cos(angle), speed * math.sin(angle), 0.0)
        # cycle settings (tri_mesh/blocks.cpp)
        self.ncycle = im.geti("ncycle", 1, prefix="")
        # steady (dtinv 0): each of the ntstep steps runs ncycle cycles again
        self.ntstep = max(im.geti("ntstep", 1, prefix=""), 1)
        self.vw = im.geti("vwcycle", 1, prefix="")
        self.itercrsn = im.geti("itercrsn", 1, prefix="")
        self.iterrfne = im.geti("iterrfne", 0, prefix="")
        self.extra = im.geti("extra_finest_levels", 0, prefix="")
        self.ngrid = im.geti("ngrid", 1, prefix="")
        self.rtol = im.getf("relative_tolerance", -1.0, prefix="")
        self.atol = im.getf("absolute_tolerance", -1.0, prefix="")
        if pairs and self.ngrid > 1:
            raise ValueError("ngrid > 1 on a periodic mesh: h-multigrid levels of periodic meshes are not supported")
        self.types = [types.get(b, 0) for b in range(m.nebd)]
        self.curved = curved
        self.log2p = im.geti("log2p", 0, prefix)
        self.grids = []

    def build(self):
        """The GPU contexts (the fine mesh and ngrid - 1 coarser meshes), boundaries set
        and the initial state loaded."""
        m, im = self.mesh, self.inmap
        self.fine = G = TriHpIns.from_input_map(m, im, self.prefix)
        if self.pairs:
            G.set_periodic(self.pairs)
        G.set_edge_flux(self.types, self.ibc)
        p = 1 << G.log2pmax
        if self.curved and p >= 2:
            from . import curved as CV
            sides, crds = [], []
            for b, key in self.curved:
                sb = m.ebdry_seg[m.ebdry_start[b]:m.ebdry_start[b + 1]]
                sides.append(sb)
                crds.append(CV.side_modes(m, sb, _level_set_projector(im, key), p))
            G.set_curved(np.concatenate(sides), np.concatenate(crds))
        self.grids = [G]
        if self.ngrid > 1:
            meshes, links = MG.mesh_hierarchy(m, self.ngrid)
            self.grids = MG.build_stack(TriHpIns, G, meshes, links, **G.params)
            for c in self.grids[1:]:
                c.set_edge_flux(self.types, self.ibc)
        G.set_state(*self.initial_state())
        return self

    def initial_state(self):
        """ibc_ins::freestream at time zero (with its x (1 - x) perturbation of u): vertex
        values, side modes of the perturbation along every side, interior modes zero (the
        quadratic perturbation lies in the vertex and side modes)."""
        from . import curved as CV
        m = self.mesh
        p = 1 << self.log2p
        sm0, im0 = p - 1, (p - 1) * (p - 2) // 2
        u0, v0, p0 = self.ibc
        amp = self.amp

        def f(x):
            return np.stack([u0 + amp * x[:, 0] * (1.0 - x[:, 0]), np.full(len(x), v0), np.full(len(x), p0)], axis=1)

        v = f(m.pnts)
        s = np.zeros((m.nseg, sm0, 3))
        if p >= 2 and amp != 0.0:
            s[:, :, :] = CV.side_data_modes(m, np.arange(m.nseg), np.zeros((m.nseg, p - 1, 2)), f, p)
        i = np.zeros((m.ntri, im0, 3))
        return v, s, i

    def run(self, ncycle: int | None = None, ntstep: int | None = None, log=None):
        """The deck's steps of cycles (tri_mesh/blocks.cpp:1160-1230); returns the max
        residual after each cycle.  A step stops at the deck's relative (to its first
        cycle) or absolute tolerance."""
        hist = []
        for step in range(self.ntstep if ntstep is None else ntstep):
            r0 = None
            for k in range(self.ncycle if ncycle is None else ncycle):
                e = MG.cycle(self.grids, self.extra, self.vw, self.itercrsn, self.iterrfne)
                hist.append(e)
                r0 = e if r0 is None else r0
                if log:
                    log(f"step {step + 1} cycle {k + 1}: max residual {e:.6e}")
                if (self.rtol > 0 and e < self.rtol * r0) or (self.atol > 0 and e < self.atol):
                    break
        return hist

    def close(self):
        for g in self.grids[::-1]:
            g.close()
