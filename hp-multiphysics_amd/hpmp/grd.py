"""The reference's text file formats for the tri_hp path.

* ``.grd`` triangle meshes, read as tri_mesh::input does for the ``grid`` type
  (tri_mesh/input.cpp:504-603): a header ``npnt: N nseg: M ntri: T`` (older files
  say ``nvrtx`` / ``nside``), then one ``index: values`` line per vertex (x y), side
  (two vertices) and triangle (three vertices), then ``nsbd:`` edge boundaries, each
  ``idnum:`` / ``number:`` followed by one line per boundary side whose first value
  is the side (the rest of the line is ignored), and ``nvbd:`` vertex boundaries,
  each ``idnum:`` / ``point:``.  Side/triangle connectivity follows
  tri_mesh::createsegtri (tri_mesh/setup.cpp:76-137): side j of a triangle runs from
  vertex (j+1)%3 to vertex (j+2)%3 and its sign is +1 when the stored side has that
  direction.
* the ``text`` restart of a tri_hp solution, written by tri_hp::output
  (tri_hp/output.cpp:136-182) and read by tri_hp::input (tri_hp/output.cpp:1395-1475):
  ``p0 = P``, a counts line, ``END OF HEADER``, then one line per vertex, per
  (side, mode) and per (triangle, interior mode), NV values each in ``%.8e`` with a
  tab after every value.  Reading a file written at another order keeps the common
  modes, zero-fills the missing ones and skips the extra ones, as the reference does.
"""
from __future__ import annotations

import numpy as np

from .mesh import TriMesh, check_mesh


def _records(lines, start, count, what):
    """`count` lines of 'index: values' from lines[start:], as lists of tokens."""
    out = []
    for k in range(count):
        if start + k >= len(lines):
            raise ValueError(f"grd file ends inside the {what} records")
        line = lines[start + k]
        if ":" not in line:
            raise ValueError(f"grd {what} record {k} has no ':' ({line!r})")
        out.append(line.split(":", 1)[1].split())
    return out, start + count


def _header_value(line, key):
    if ":" not in line:
        raise ValueError(f"grd: expected '{key}:' line, got {line!r}")
    return int(line.split(":", 1)[1].split()[0])


def read_grd(path: str, essential=0b011) -> TriMesh:
    """Read a ``.grd`` mesh into a TriMesh.  ``essential`` is the essential-variable
    bit mask of every edge boundary (an int), or a dict {idnum: mask} (missing ids
    get 0, natural).  Vertex boundaries are returned in ``mesh.vbdry`` as
    (idnum, vertex) pairs."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln.strip()]
    head = lines[0].replace(":", " : ").split()
    nums = [int(head[i + 2]) for i in range(len(head)) if head[i] in ("npnt", "nvrtx", "nseg", "nside", "ntri")]
    if len(nums) != 3:
        raise ValueError(f"grd header not understood: {lines[0]!r}")
    npnt, nseg, ntri = nums
    pos = 1
    rec, pos = _records(lines, pos, npnt, "vertex")
    pnts = np.array([[float(r[0]), float(r[1])] for r in rec], dtype=np.float64).reshape(npnt, 2)
    rec, pos = _records(lines, pos, nseg, "side")
    seg_pnt = np.array([[int(r[0]), int(r[1])] for r in rec], dtype=np.int64).reshape(nseg, 2)
    rec, pos = _records(lines, pos, ntri, "triangle")
    tri_pnt = np.array([[int(r[0]), int(r[1]), int(r[2])] for r in rec], dtype=np.int64).reshape(ntri, 3)
    if seg_pnt.size and (seg_pnt.min() < 0 or seg_pnt.max() >= npnt):
        raise ValueError("grd side refers to a missing vertex")
    if tri_pnt.size and (tri_pnt.min() < 0 or tri_pnt.max() >= npnt):
        raise ValueError("grd triangle refers to a missing vertex")

    # createsegtri: match each triangle side to its stored side
    lookup = {}
    for s, (a, b) in enumerate(seg_pnt):
        lookup[(min(a, b), max(a, b))] = s
    tri_seg = np.empty((ntri, 3), dtype=np.int64)
    tri_sgn = np.empty((ntri, 3), dtype=np.int64)
    for t in range(ntri):
        for j in range(3):
            a, b = tri_pnt[t, (j + 1) % 3], tri_pnt[t, (j + 2) % 3]
            s = lookup.get((min(a, b), max(a, b)))
            if s is None:
                raise ValueError(f"didn't match seg: triangle {t} side {j} ({a}, {b})")
            tri_seg[t, j] = s
            tri_sgn[t, j] = 1 if seg_pnt[s, 0] == a else -1

    nebd = _header_value(lines[pos], "nsbd")
    pos += 1
    starts, segs, ids = [0], [], []
    for _ in range(nebd):
        ids.append(_header_value(lines[pos], "idnum"))
        n = _header_value(lines[pos + 1], "number")
        rec, pos = _records(lines, pos + 2, n, "boundary side")
        segs.extend(int(r[0]) for r in rec)
        starts.append(len(segs))
    vbdry = []
    if pos < len(lines):
        nvbd = _header_value(lines[pos], "nvbd")
        pos += 1
        for _ in range(nvbd):
            vid = _header_value(lines[pos], "idnum")
            vbdry.append((vid, _header_value(lines[pos + 1], "point")))
            pos += 2
    if isinstance(essential, dict):
        ess = [int(essential.get(i, 0)) for i in ids]
    else:
        ess = [int(essential)] * nebd
    m = TriMesh(
        pnts=pnts,
        seg_pnt=seg_pnt.astype(np.int32),
        tri_pnt=tri_pnt.astype(np.int32),
        tri_seg=tri_seg.astype(np.int32),
        tri_sgn=tri_sgn.astype(np.int32),
        ebdry_start=np.array(starts, dtype=np.int32),
        ebdry_seg=np.array(segs, dtype=np.int32),
        ebdry_essential=np.array(ess, dtype=np.int32),
        ebdry_names=[f"b{i}" for i in ids],
        ebdry_ids=ids,
        vbdry=vbdry,
    )
    check_mesh(m)
    return m


def write_grd(path: str, m: TriMesh) -> None:
    """Write ``m`` in the ``.grd`` layout read_grd (and tri_mesh::input) reads."""
    ids = m.ebdry_ids or list(range(1, m.nebd + 1))
    with open(path, "w") as f:
        f.write(f"npnt: {m.npnt} nseg: {m.nseg} ntri: {m.ntri}\n")
        for i, (x, y) in enumerate(m.pnts):
            f.write(f"{i}: {x:.10e} {y:.10e}\n")
        for i, (a, b) in enumerate(m.seg_pnt):
            f.write(f"{i}: {a} {b}\n")
        for i, (a, b, c) in enumerate(m.tri_pnt):
            f.write(f"{i}: {a} {b} {c}\n")
        f.write(f"nsbd: {m.nebd}\n")
        for k in range(m.nebd):
            s = m.ebdry_seg[m.ebdry_start[k]:m.ebdry_start[k + 1]]
            f.write(f"idnum: {ids[k]}\nnumber: {len(s)}\n")
            for j, sj in enumerate(s):
                f.write(f"{j}: {sj}\n")
        vb = m.vbdry
        f.write(f"nvbd: {len(vb)}\n")
        for vid, v in vb:
            f.write(f"idnum: {vid}\npoint: {v}\n")


def write_restart_text(path: str, p: int, v, s, i) -> None:
    """tri_hp::output, ``text`` type: v (npnt, NV), s (nseg, p-1, NV), i (ntri, im, NV)."""
    v, s, i = (np.asarray(a, dtype=np.float64) for a in (v, s, i))

    def row(a):
        return "".join(f"{x:.8e}\t" for x in a) + "\n"

    with open(path, "w") as f:
        f.write(f"p0 = {p}\n")
        f.write(f"npnt = {v.shape[0]}, nseg = {s.shape[0]}, ntri = {i.shape[0]}\n")
        f.write("END OF HEADER\n")
        for a in v:
            f.write(row(a))
        for sg in s:
            for a in sg[: p - 1]:
                f.write(row(a))
        for tr in i:
            for a in tr[: (p - 1) * (p - 2) // 2]:
                f.write(row(a))


def read_restart_text(path: str, p0: int, npnt: int, nseg: int, ntri: int, nv: int):
    """tri_hp::input, ``text`` type, into arrays at order p0 (side stride p0-1,
    interior stride (p0-1)(p0-2)/2).  Returns (v, s, i)."""
    with open(path) as f:
        first = f.readline()
        pin = int(first.split("=", 1)[1].split()[0])
        for line in f:
            if line.startswith("E"):
                break
        tok = np.array(f.read().split(), dtype=np.float64)
    pmin = min(p0, pin)
    im0 = (p0 - 1) * (p0 - 2) // 2
    v = np.zeros((npnt, nv))
    s = np.zeros((nseg, max(p0 - 1, 0), nv))
    i = np.zeros((ntri, max(im0, 0), nv))
    k = 0

    def take(n):
        nonlocal k
        if k + n > tok.size:
            raise ValueError("restart file ends early")
        out = tok[k:k + n]
        k += n
        return out

    v[:] = take(npnt * nv).reshape(npnt, nv)
    for sd in range(nseg):
        if pmin - 1 > 0:
            s[sd, :pmin - 1] = take((pmin - 1) * nv).reshape(pmin - 1, nv)
        take(max(pin - p0, 0) * nv)
    for t in range(ntri):
        idx = 0
        for m in range(1, pmin - 1):
            n = pmin - 1 - m
            i[t, idx:idx + n] = take(n * nv).reshape(n, nv)
            idx += n + (p0 - pmin)          # the modes pin lacks stay zero
            take(max(pin - p0, 0) * nv)     # the modes p0 lacks are skipped
        for m in range(max(pmin - 1, 1), pin - 1):
            take((pin - 1 - m) * nv)
    return v, s, i
