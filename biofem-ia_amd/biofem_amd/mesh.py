"""Mesh inputs of the force path: triangle topology, ghost layer, Loop patch table.

The reference reads meshes from nlist/elist/blist files or builds the LV meshes in
lv_geometry*.c, then derives the patch table (src/bending.c IrrVer/Patch), the ghost layer
(GlobalGhostInit) and the reference areas (src/io.c AreaNormal).  Here the synthetic
meshes are generated directly and the derived tables come from the library's native
topology builders (bf_topo_*), which restate those routines with hashed edge lookups.

Flat membranes use a structured lattice with one full ghost ring (corners included), so
that every real node has valence 6 in the extended triangulation and every element is a
regular 12-node Loop patch.  Ghost positions follow the reference mirror rule
ghost = x[b] + x[a] - x[c] (src/bending.c:1731-1733): across each boundary edge (a, b)
from the opposite vertex c; the four lattice corners and the two row/column ends use the
point reflection 2 x[a] - x[c] (a = b), which equals the lattice position for an
undeformed grid.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


@dataclass
class Mesh:
    n_v: int
    n_nodes: int
    n_elmt: int
    x0: np.ndarray
    y0: np.ndarray
    z0: np.ndarray
    tri: np.ndarray            # [n_tri, 3] int32: real elements first, then ghost triangles
    patch: np.ndarray          # [n_elmt, stride] int32
    ire: np.ndarray
    irv: np.ndarray
    val: np.ndarray
    dA0: np.ndarray
    n_fib: np.ndarray          # [n_elmt, 3]
    gamma_scale: np.ndarray
    ghost_rules: np.ndarray    # [n_ghost, 3] int32 (b, a, c): ghost = x[b] + x[a] - x[c]
    activation_delay: np.ndarray | None = None
    n_elmt_base: int = 0
    name: str = "mesh"
    extra: dict = field(default_factory=dict)

    @property
    def stride(self) -> int:
        return int(self.patch.shape[1])

    def nen(self) -> np.ndarray:
        return np.where(self.ire == 0, 12, self.val + 6).astype(np.int32)

    def apply_ghosts(self, x, y, z) -> None:
        """Host-side GlobalGhost on numpy arrays (mesh generation only)."""
        g = self.ghost_rules
        if len(g) == 0:
            return
        idx = self.n_v + np.arange(len(g))
        for a in (x, y, z):
            a[idx] = a[g[:, 0]] + a[g[:, 1]] - a[g[:, 2]]

    def footprint_bytes(self) -> int:
        """Device bytes the library holds for this mesh (see bf_mesh_upload)."""
        E, N, S = self.n_elmt, self.n_nodes, max(12, int(self.nen().max()) if self.n_elmt else 12)
        per_e = S * 4 + 4 + 15 * 8 + 8 + 24 + 8 + 3 * S * 8 + 5 * 8 + 12 * 4
        per_n = 4 * 8 + 3 * 8 + 4
        return E * per_e + N * per_n

    def desc(self):
        """bf_mesh_desc over this mesh's arrays (the arrays must outlive the call)."""
        self._keep = [np.ascontiguousarray(self.patch, dtype=np.int32),
                      np.ascontiguousarray(self.n_fib, dtype=np.float64),
                      np.ascontiguousarray(self.ghost_rules, dtype=np.int32),
                      np.ascontiguousarray(self.tri[: self.n_elmt, 0], dtype=np.int32),
                      np.ascontiguousarray(self.tri[: self.n_elmt, 1], dtype=np.int32),
                      np.ascontiguousarray(self.tri[: self.n_elmt, 2], dtype=np.int32)]
        d = L.BfMeshDesc()
        d.n_v, d.n_nodes, d.n_elmt = self.n_v, self.n_nodes, self.n_elmt
        d.n_elmt_base = self.n_elmt_base
        d.patch_stride = self.stride
        d.patch = L.ip(self._keep[0])
        d.ire, d.val = L.ip(self.ire), L.ip(self.val)
        d.nv1, d.nv2, d.nv3 = L.ip(self._keep[3]), L.ip(self._keep[4]), L.ip(self._keep[5])
        d.x0, d.y0, d.z0 = L.dp(self.x0), L.dp(self.y0), L.dp(self.z0)
        d.dA0 = L.dp(self.dA0)
        d.n_fib = L.dp(self._keep[1])
        d.gamma_scale = L.dp(self.gamma_scale)
        d.activation_delay = L.dp(self.activation_delay) if self.activation_delay is not None else None
        d.n_ghost_rules = len(self.ghost_rules)
        d.ghost_rules = L.ip(self._keep[2]) if len(self.ghost_rules) else None
        return d


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def build_patch(n_v, n_nodes, n_elmt, tri, ire, irv, val, stride=16, n_elmt_base=0) -> np.ndarray:
    lib = L.lib()
    tri = _i32(tri)
    nv1, nv2, nv3 = _i32(tri[:, 0]), _i32(tri[:, 1]), _i32(tri[:, 2])
    patch = np.full((n_elmt, stride), L.SENTINEL, dtype=np.int32)
    L.check("bf_topo_patch", lib.bf_topo_patch(n_v, n_nodes, n_elmt, n_elmt_base, len(tri),
                                                L.ip(nv1), L.ip(nv2), L.ip(nv3), L.ip(ire),
                                                L.ip(irv), L.ip(val), stride, L.ip(patch)))
    return patch


def area0(tri, x, y, z) -> np.ndarray:
    lib = L.lib()
    tri = _i32(tri)
    nv1, nv2, nv3 = _i32(tri[:, 0]), _i32(tri[:, 1]), _i32(tri[:, 2])
    dA = np.zeros(len(tri))
    L.check("bf_topo_area0", lib.bf_topo_area0(len(tri), L.ip(nv1), L.ip(nv2), L.ip(nv3),
                                                L.dp(x), L.dp(y), L.dp(z), L.dp(dA)))
    return dA


def flat_membrane(nx: int, ny: int, Lx: float = 1.0, Ly: float = 0.75, jitter: float = 1e-3,
                  seed: int = 0, stride: int = 16) -> Mesh:
    """Rectangle Lx x Ly split into 2 nx ny triangles ('/' diagonals), one ghost ring.

    Interior nodes are displaced by U(-1,1) * jitter * h in x and y, h = Lx / nx, with
    numpy default_rng(seed): one uniform draw per real node for x, then one per real node
    for y (boundary draws are discarded).
    """
    hx, hy = Lx / nx, Ly / ny
    nvx = nx + 1
    n_v = nvx * (ny + 1)
    I, J = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    I, J = I.ravel(), J.ravel()
    x = I * hx
    y = J * hy
    rng = np.random.default_rng(seed)
    dx = rng.uniform(-1.0, 1.0, n_v)
    dy = rng.uniform(-1.0, 1.0, n_v)
    inner = (I > 0) & (I < nx) & (J > 0) & (J < ny)
    x = x + np.where(inner, dx * jitter * hx, 0.0)
    y = y + np.where(inner, dy * jitter * hx, 0.0)

    def nid(i, j):
        return j * nvx + i

    # ghost nodes and their mirror rules (b, a, c): ghost = x[b] + x[a] - x[c]
    gpos, rules = [], []
    for i in range(-1, nx + 1):          # bottom row j = -1
        gpos.append((i, -1))
        if 0 <= i <= nx - 1:
            rules.append((nid(i, 0), nid(i + 1, 0), nid(i + 1, 1)))
        elif i == -1:
            rules.append((nid(0, 0), nid(0, 0), nid(1, 1)))
        else:
            rules.append((nid(nx, 0), nid(nx, 0), nid(nx, 1)))
    for i in range(0, nx + 2):           # top row j = ny + 1
        gpos.append((i, ny + 1))
        if 1 <= i <= nx:
            rules.append((nid(i - 1, ny), nid(i, ny), nid(i - 1, ny - 1)))
        elif i == 0:
            rules.append((nid(0, ny), nid(0, ny), nid(0, ny - 1)))
        else:
            rules.append((nid(nx, ny), nid(nx, ny), nid(nx - 1, ny - 1)))
    for j in range(0, ny + 1):           # left column i = -1
        gpos.append((-1, j))
        if j <= ny - 1:
            rules.append((nid(0, j), nid(0, j + 1), nid(1, j + 1)))
        else:
            rules.append((nid(0, ny), nid(0, ny), nid(1, ny)))
    for j in range(0, ny + 1):           # right column i = nx + 1
        gpos.append((nx + 1, j))
        if j >= 1:
            rules.append((nid(nx, j - 1), nid(nx, j), nid(nx - 1, j - 1)))
        else:
            rules.append((nid(nx, 0), nid(nx, 0), nid(nx - 1, 0)))
    gid = {p: n_v + k for k, p in enumerate(gpos)}
    n_g = len(gpos)
    n_nodes = n_v + n_g

    def node(i, j):
        if 0 <= i <= nx and 0 <= j <= ny:
            return nid(i, j)
        return gid.get((i, j), -1)

    # real triangles, cell-major
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny))
    ci, cj = ci.ravel(), cj.ravel()
    a = nid(ci, cj)
    b = nid(ci + 1, cj)
    c = nid(ci + 1, cj + 1)
    d = nid(ci, cj + 1)
    t1 = np.stack([a, b, c], 1)
    t2 = np.stack([a, c, d], 1)
    real = np.stack([t1, t2], 1).reshape(-1, 3)
    # ghost triangles: lattice cells of the ghost ring whose corners all exist
    gt = []
    for j in range(-1, ny + 1):
        for i in range(-1, nx + 1):
            if 0 <= i < nx and 0 <= j < ny:
                continue
            p00, p10, p11, p01 = node(i, j), node(i + 1, j), node(i + 1, j + 1), node(i, j + 1)
            if min(p00, p10, p11) >= 0:
                gt.append((p00, p10, p11))
            if min(p00, p11, p01) >= 0:
                gt.append((p00, p11, p01))
    tri = np.concatenate([real, np.array(gt, dtype=np.int64).reshape(-1, 3)]).astype(np.int32)
    E = len(real)

    x0 = np.zeros(n_nodes); y0 = np.zeros(n_nodes); z0 = np.zeros(n_nodes)
    x0[:n_v], y0[:n_v] = x, y
    rules = np.array(rules, dtype=np.int32)
    m = Mesh(n_v=n_v, n_nodes=n_nodes, n_elmt=E, x0=x0, y0=y0, z0=z0, tri=tri,
             patch=np.zeros((E, stride), np.int32), ire=np.zeros(E, np.int32),
             irv=np.zeros(E, np.int32), val=np.full(E, 6, np.int32), dA0=np.zeros(E),
             n_fib=np.tile([1.0, 0.0, 0.0], (E, 1)), gamma_scale=np.ones(E),
             ghost_rules=rules, name=f"flat{nx}x{ny}")
    m.apply_ghosts(m.x0, m.y0, m.z0)
    m.patch = build_patch(n_v, n_nodes, E, tri, m.ire, m.irv, m.val, stride)
    if (m.patch[:, :12] == L.SENTINEL).any():
        raise RuntimeError("flat lattice produced an incomplete patch")
    m.dA0 = area0(tri[:E], m.x0, m.y0, m.z0)
    m.extra.update(nx=nx, ny=ny, Lx=Lx, Ly=Ly)
    return m


def centroids(m: Mesh, x=None, y=None, z=None) -> np.ndarray:
    x = m.x0 if x is None else x
    y = m.y0 if y is None else y
    z = m.z0 if z is None else z
    t = m.tri[: m.n_elmt]
    return np.stack([(x[t[:, 0]] + x[t[:, 1]] + x[t[:, 2]]) / 3.0,
                     (y[t[:, 0]] + y[t[:, 1]] + y[t[:, 2]]) / 3.0,
                     (z[t[:, 0]] + z[t[:, 1]] + z[t[:, 2]]) / 3.0], 1)
