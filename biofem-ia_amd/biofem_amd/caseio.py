"""The reference's case files, VTK frames and state checkpoints (include/biofem_io.h).

read_case loads a BioFEM case directory (nlist/elist/blist, src/io.c:34-537) and runs the
reference's setup on it -- GlobalGhostInit over the boundary rings, IrrVer, Patch,
AreaNormal -- natively; write_case writes a mesh back in the same layout; write_vtk writes
the reference's surface frame (src/io.c:540-930); write_state / read_state the x, xn, xd,
xdd checkpoints (src/io.c:1272-1376, 1511-1640).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib as L
from .mesh import Mesh, area0, build_patch

STATE_RAW = 0
STATE_VECVIEW = 1
_TAGS = ("x", "xn", "xd", "xdd")


def _b(path: str) -> bytes:
    return os.fsencode(path)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def dimension(in_dir: str, ibi: int = 0) -> dict:
    """Dimension (src/io.c:34-83): n_v, n_elmt, n_edge, n_ghosts, and the summed length of
    the boundary lists."""
    lib = L.lib()
    v = [C.c_int() for _ in range(5)]
    L.check("bf_io_dimension", lib.bf_io_dimension(_b(in_dir), ibi, *[C.byref(a) for a in v]))
    return dict(zip(("n_v", "n_elmt", "n_edge", "n_ghosts", "sum_n_bnodes"), (a.value for a in v)))


def read_lists(in_dir: str, ibi: int = 0, char_length=(1.0, 1.0, 1.0)) -> dict:
    """Input (src/io.c:333-537): coordinates divided by char_length, 0-based corners and
    boundary lists, and the neighbour vertices nv4/nv5/nv6 (1000000: none)."""
    lib = L.lib()
    d = dimension(in_dir, ibi)
    n_v, E, ne, sb = d["n_v"], d["n_elmt"], d["n_edge"], d["sum_n_bnodes"]
    x, y, z = (np.zeros(n_v) for _ in range(3))
    nv = [np.zeros(E, np.int32) for _ in range(6)]
    nb = np.zeros(max(ne, 1), np.int32)
    bn = np.zeros(max(sb, 1), np.int32)
    cl = _f64(char_length)
    L.check("bf_io_input", lib.bf_io_input(_b(in_dir), ibi, L.dp(cl), L.dp(x), L.dp(y), L.dp(z),
                                           *[L.ip(a) for a in nv[:3]], L.ip(nb), L.ip(bn),
                                           *[L.ip(a) for a in nv[3:]]))
    out = dict(d)
    out.update(x=x, y=y, z=z, tri=np.stack(nv[:3], 1), nv456=np.stack(nv[3:], 1),
               n_bnodes=nb[:ne].copy(), bnodes=bn[:sb].copy())
    return out


def _rings(n_bnodes, bnodes):
    """The boundary lists as closed rings.  Lists that chain end to start (the four open
    sides of a rectangle written by `biofem sweep`, bin/biofem:533-548) are joined into one
    ring with the shared corners once; GlobalGhostInit walks every list as a closed ring
    (src/bending.c:1689-1755), which open sides are not."""
    lists, k = [], 0
    for n in n_bnodes:
        lists.append([int(v) for v in bnodes[k:k + n]])
        k += n
    if len(lists) > 1 and all(lists[i][-1] == lists[(i + 1) % len(lists)][0] for i in range(len(lists))):
        ring = []
        for li in lists:
            ring.extend(li[:-1])
        return [ring]
    return lists


def read_case(in_dir: str, ibi: int = 0, char_length=(1.0, 1.0, 1.0), stride: int = 16) -> Mesh:
    """A BioFEM case directory as a Mesh: read_lists, then GlobalGhostInit (one ghost and
    two ghost triangles per boundary edge, mirror rule x[b] + x[a] - x[c]), IrrVer, Patch
    and AreaNormal of the reference state.  Fibres are e_x and gamma_scale 1 until the
    caller sets its case's."""
    lib = L.lib()
    r = read_lists(in_dir, ibi, char_length)
    n_v, E = r["n_v"], r["n_elmt"]
    rings = _rings(r["n_bnodes"], r["bnodes"])
    ring_len = np.array([len(g) for g in rings], np.int32)
    ring_nodes = np.array([v for g in rings for v in g], np.int32)
    G = int(ring_len.sum())
    T = r["tri"]
    nv = [np.zeros(E + 2 * G, np.int32) for _ in range(3)]
    for k in range(3):
        nv[k][:E] = T[:, k]
    rules = np.zeros(max(3 * G, 1), np.int32)
    if G:
        L.check("bf_topo_ghost_ring", lib.bf_topo_ghost_ring(
            n_v, E, 0, len(rings), L.ip(ring_len), L.ip(ring_nodes), L.ip(nv[0]), L.ip(nv[1]),
            L.ip(nv[2]), L.ip(rules)))
    ire, irv, val = (np.zeros(E, np.int32) for _ in range(3))
    L.check("bf_topo_irrver", lib.bf_topo_irrver(
        n_v, E, 0, L.ip(_i32(T[:, 0])), L.ip(_i32(T[:, 1])), L.ip(_i32(T[:, 2])), G,
        L.ip(ring_nodes if G else np.zeros(1, np.int32)), L.ip(ire), L.ip(irv), L.ip(val)))
    n_nodes = n_v + G
    x0, y0, z0 = (np.zeros(n_nodes) for _ in range(3))
    x0[:n_v], y0[:n_v], z0[:n_v] = r["x"], r["y"], r["z"]
    tri = np.stack(nv, 1).astype(np.int32)
    m = Mesh(n_v=n_v, n_nodes=n_nodes, n_elmt=E, x0=x0, y0=y0, z0=z0, tri=tri,
             patch=np.zeros((E, stride), np.int32), ire=ire, irv=irv, val=val, dA0=np.zeros(E),
             n_fib=np.tile([1.0, 0.0, 0.0], (E, 1)), gamma_scale=np.ones(E),
             ghost_rules=rules[: 3 * G].reshape(-1, 3), name=f"case{ibi:02d}")
    m.apply_ghosts(m.x0, m.y0, m.z0)
    m.patch = build_patch(n_v, n_nodes, E, tri, ire, irv, val, stride)
    m.dA0 = area0(tri[:E], m.x0, m.y0, m.z0)
    m.extra.update(rings=rings, blist_n_ghosts=r["n_ghosts"], nv456=r["nv456"])
    return m


def boundary_rings(tri: np.ndarray) -> list:
    """Closed boundary rings of a triangle mesh (edges of one triangle), each walked in the
    triangles' winding direction -- the order GlobalGhostInit needs for ghost triangles
    wound like their neighbours."""
    tri = np.asarray(tri)
    count = {}
    for a, b, c in tri:
        for u, v in ((a, b), (b, c), (c, a)):
            key = (min(u, v), max(u, v))
            count[key] = count.get(key, 0) + 1
    nxt = {}
    for a, b, c in tri:
        for u, v in ((a, b), (b, c), (c, a)):
            if count[(min(u, v), max(u, v))] == 1:
                nxt[int(u)] = int(v)
    rings, seen = [], set()
    for s in sorted(nxt):
        if s in seen:
            continue
        ring, v = [], s
        while v not in seen:
            seen.add(v)
            ring.append(v)
            v = nxt[v]
        rings.append(ring)
    return rings


def write_case(out_dir: str, m: Mesh, ibi: int = 0, rings: list | None = None) -> None:
    """nlist/elist/blist of the real nodes and triangles of m (rings: 0-based closed
    boundary lists; default boundary_rings of the real triangles)."""
    lib = L.lib()
    os.makedirs(out_dir, exist_ok=True)
    E, n_v = m.n_elmt, m.n_v
    T = np.asarray(m.tri[:E])
    rings = boundary_rings(T) if rings is None else rings
    nb = np.array([len(g) for g in rings], np.int32)
    bn = np.array([v for g in rings for v in g], np.int32)
    L.check("bf_io_write_lists", lib.bf_io_write_lists(
        _b(out_dir), ibi, n_v, L.dp(_f64(m.x0[:n_v])), L.dp(_f64(m.y0[:n_v])), L.dp(_f64(m.z0[:n_v])),
        E, L.ip(_i32(T[:, 0])), L.ip(_i32(T[:, 1])), L.ip(_i32(T[:, 2])), len(rings),
        L.ip(nb if len(nb) else np.zeros(1, np.int32)), L.ip(bn if len(bn) else np.zeros(1, np.int32)),
        int(nb.sum())))


def write_vtk(out_dir: str, m: Mesh, ti: int, x, y, z, ibi: int = 0, fint=None, fext=None,
              fdyn=None, fcnt=None, u=None, xdd=None, mass=None, ie=None, contact=None,
              dA=None, gamma=None, muscle_activation: bool = False, manufactured: bool = False,
              lv_geom_process: bool = False) -> str:
    """Output's surface frame (src/io.c:540-930) of state (x, y, z); vectors interleaved
    [3 n_nodes] as the reference's Vecs.  Returns the file path."""
    lib = L.lib()
    keep = []

    def d(a):
        if a is None:
            return None
        a = _f64(a)
        keep.append(a)
        return L.dp(a)

    def i(a):
        if a is None:
            return None
        a = _i32(a)
        keep.append(a)
        return L.ip(a)

    E = m.n_elmt
    T = np.asarray(m.tri[:E])
    fr = L.BfVtkFrame()
    fr.n_v, fr.n_nodes, fr.n_elmt = m.n_v, m.n_nodes, E
    fr.x, fr.y, fr.z = d(x), d(y), d(z)
    fr.x0, fr.y0, fr.z0 = d(m.x0), d(m.y0), d(m.z0)
    fr.nv1, fr.nv2, fr.nv3 = i(T[:, 0]), i(T[:, 1]), i(T[:, 2])
    fr.contact, fr.m, fr.IE = i(contact), d(mass), d(ie)
    fr.Fint, fr.Fext, fr.Fdyn, fr.Fcnt, fr.u, fr.xdd = d(fint), d(fext), d(fdyn), d(fcnt), d(u), d(xdd)
    fr.dA, fr.dA0 = d(dA), d(m.dA0)
    fr.n_fib = d(np.asarray(m.n_fib).reshape(-1))
    fr.gamma = d(gamma)
    fr.muscle_activation, fr.manufactured = int(muscle_activation), int(manufactured)
    fr.lv_geom_process = int(lv_geom_process)
    L.check("bf_io_write_vtk", lib.bf_io_write_vtk(_b(out_dir), ibi, ti, C.byref(fr)))
    return os.path.join(out_dir or ".", f"surface{ibi:02d}_{ti:05d}.vtk")


def write_state(dir_: str, ti: int, n_total: int, x=None, xn=None, xd=None, xdd=None,
                ibi: int = 0, layout: int = STATE_RAW) -> None:
    """LocationOut (src/io.c:1272-1376): the given vectors [3 n_total], big-endian."""
    lib = L.lib()
    arrs = [None if a is None else _f64(a) for a in (x, xn, xd, xdd)]
    for a in arrs:
        if a is not None and a.size != 3 * n_total:
            raise ValueError(f"state vector of {a.size} values, expected {3 * n_total}")
    L.check("bf_io_write_state", lib.bf_io_write_state(
        _b(dir_), ibi, ti, n_total, *[None if a is None else L.dp(a) for a in arrs], layout))


def read_state(dir_: str, ti: int, n_total: int, ibi: int = 0, tags=_TAGS):
    """LocationIn (src/io.c:1511-1640): {tag: [3 n_total]} for the requested tags, and the
    layout found (STATE_RAW or STATE_VECVIEW)."""
    lib = L.lib()
    out = {t: np.zeros(3 * n_total) for t in tags}
    lay = C.c_int(-1)
    L.check("bf_io_read_state", lib.bf_io_read_state(
        _b(dir_), ibi, ti, n_total, *[L.dp(out[t]) if t in out else None for t in _TAGS],
        C.byref(lay)))
    return out, lay.value
