"""h- and p-multigrid over a stack of meshes: the multigrid part of ``block``
(tri_mesh/blocks.cpp).

``grids[0]`` is the fine mesh at its full order; ``grids[1:]`` are coarser meshes at
p = 1, each linked to the one above it (``link_fine``).  The first
``extra_finest_levels`` multigrid levels lower p on the fine mesh, the following ones
step down the mesh stack, exactly as block::cycle maps a level to a grid
(tri_mesh/blocks.cpp:1274-1275).  The cycle works on any object with the tri_hp
operations (setup_preconditioner, update, rsdl, maxres, mg_restrict, mg_prolongate),
the GPU contexts of ``hpmp.tri_hp`` as well as CPU restatements of them.
"""
from __future__ import annotations

from . import mesh as M


def fused_available(grids) -> bool:
    """True when every grid is a GPU tri context (hpg_mg_cycle runs the whole cycle)."""
    from .tri_hp import TriHpIns
    return all(isinstance(g, TriHpIns) for g in grids)


def cycle(grids, extra_finest_levels: int, vw: int = 1, itercrsn: int = 1, iterrfne: int = 0,
          lvl: int = 0, mglvls: int | None = None, evaluate_preconditioner: bool = True,
          fused: bool | None = None) -> float:
    """block::cycle (tri_mesh/blocks.cpp:1267-1348) without the error-control level.

    Returns the fine-level max residual measured after the first smoothing pass
    (``maxres`` on the entry level), like the reference; raises FloatingPointError on a
    preconditioner error or a residual that is NaN or exactly zero (``throw 1``).

    On GPU contexts (``fused`` None or True) the whole cycle is one call of
    hpg_mg_cycle: the same kernels in the same order, with no host round trip inside the
    cycle, so setup_preconditioner errors are reported after the cycle rather than at
    the failing level.  ``fused=False`` runs the level recursion call by call."""
    if lvl == 0 and mglvls is None and evaluate_preconditioner and fused is not False and fused_available(grids):
        return grids[0].mg_cycle(grids, extra_finest_levels, vw, itercrsn, iterrfne)
    if fused:
        raise ValueError("the fused cycle needs GPU tri contexts, the entry level and evaluate_preconditioner")
    ngrid = len(grids)
    if mglvls is None:
        mglvls = ngrid + extra_finest_levels
    gl = min(max(lvl - extra_finest_levels, 0), ngrid - 1)
    glp = min(max(lvl - extra_finest_levels + 1, 0), ngrid - 1)
    g, gp = grids[gl], grids[glp]
    error = 0.0
    for vcount in range(vw):
        if evaluate_preconditioner and g.setup_preconditioner():
            raise FloatingPointError("setup_preconditioner reported an error")
        for _ in range(itercrsn):
            g.update()
        if lvl + vcount == 0:
            error = g.maxres()[0]
            if not error > 0.0:
                raise FloatingPointError("maxres is not positive (NaN or exact zero)")
        if lvl == mglvls - 1:
            return error
        g.rsdl()
        gp.mg_restrict()
        cycle(grids, extra_finest_levels, vw, itercrsn, iterrfne, lvl + 1, mglvls, evaluate_preconditioner, fused=False)
        gp.mg_prolongate()
        if evaluate_preconditioner and iterrfne:
            if g.setup_preconditioner():
                raise FloatingPointError("setup_preconditioner reported an error")
        for _ in range(iterrfne):
            g.update()
    return error


def square_hierarchy(fine_mesh, n_fine: int, ngrid: int):
    """Meshes and connection tables for ``ngrid`` levels of a structured unit-square mesh
    (each coarse mesh takes every other fine vertex).  Returns ``(meshes, links)`` where
    ``links[k] = (cc_tri, cc_wt, fc_tri, fc_wt)`` joins ``meshes[k]`` to ``meshes[k+1]``."""
    meshes, links, n = [fine_mesh], [], n_fine
    for _ in range(1, ngrid):
        c = M.coarsen_square(meshes[-1], n)
        links.append(M.connect(meshes[-1], n, c))
        meshes.append(c)
        n //= 2
    return meshes, links


def mesh_hierarchy(fine_mesh, ngrid: int, n_fine: int | None = None):
    """Meshes and connection tables of an h-multigrid stack: the structured coarsening
    of ``unit_square(n_fine)`` when n_fine is given, the unstructured one of
    hpmp.coarsen (vertex removal and Delaunay, as tri_mesh::coarsen) otherwise."""
    if n_fine is not None:
        return square_hierarchy(fine_mesh, n_fine, ngrid)
    from .coarsen import hierarchy
    return hierarchy(fine_mesh, ngrid)


def build_stack(cls, fine, meshes, links, **params):
    """Coarse levels of the same physics as ``fine`` on ``meshes[1:]`` at p = 1, linked
    downwards.  ``cls`` is the level class (GPU or CPU); ``params`` are its physical
    and relaxation parameters.  Returns the grid list with ``fine`` first."""
    grids = [fine]
    for k in range(1, len(meshes)):
        g = cls(meshes[k], log2p=0, **params)
        g.link_fine(grids[-1], *links[k - 1])
        grids.append(g)
    return grids
