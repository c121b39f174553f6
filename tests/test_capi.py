"""The C-ABI library loads without a GPU and exports every entry point the header
declares; the ctypes structures match the header's layouts."""
import ctypes as C
import os
import re

from hpmp import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hpgpu.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w ]*\**\s*\**(hpg_\w+)\s*\(", text, flags=re.M))


def test_header_lists_the_bound_entry_points():
    decl = _declared()
    assert len(decl) >= 40
    assert decl == set(_lib.EXPORTS), (decl ^ set(_lib.EXPORTS))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    missing = [s for s in sorted(_declared()) if not hasattr(L, s)]
    assert not missing, missing


def test_struct_layouts():
    # hpg_tri_mesh: 3 ints, 5 pointers, 1 int, 3 pointers (natural alignment)
    assert C.sizeof(_lib.TriMeshC) == 8 + 8 + 5 * 8 + 8 + 3 * 8
    # hpg_params: physics, log2p, 5 doubles, nstage, alpha[8], beta[8], cfl[4], gamma, prandtl, Rgas, bc[5]
    assert C.sizeof(_lib.ParamsC) == 8 + 5 * 8 + 8 + 20 * 8 + 3 * 8 + 5 * 8
    assert _lib.ParamsC.bc.offset == C.sizeof(_lib.ParamsC) - 40
