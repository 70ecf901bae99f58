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

    def footprint_bytes(self, state: bool = False) -> int:
        """Device bytes the library holds for this mesh (bf_mesh_upload: patch, meta, geom0
        cache, dA0, fibre, gamma scale, element-force buffer, C33, flags, hand-over lists,
        node CSR; coordinates and forces), plus the time-stepping state when `state`."""
        E, N, S = self.n_elmt, self.n_nodes, max(12, int(self.nen().max()) if self.n_elmt else 12)
        per_e = S * 4 + 4 + 15 * 8 + 8 + 24 + 8 + 3 * S * 8 + 5 * 8 + 1 + 12 * 4 + S * 4  # pos: S*4
        per_n = 3 * 8 + 3 * 8 + 4 + (3 * 8 + 8 + 1 if state else 0)
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


def lumped_mass(m: Mesh, rho: float, h0: float) -> np.ndarray:
    """Row-sum lumped masses of the reference shell: each real triangle gives a third of
    rho * h0 * dA0 to each of its corners (ghost nodes get 0)."""
    mass = np.zeros(m.n_nodes)
    share = rho * h0 * m.dA0 / 3.0
    for k in range(3):
        np.add.at(mass, m.tri[: m.n_elmt, k], share)
    return mass


def centroids(m: Mesh, x=None, y=None, z=None) -> np.ndarray:
    x = m.x0 if x is None else x
    y = m.y0 if y is None else y
    z = m.z0 if z is None else z
    t = m.tri[: m.n_elmt]
    return np.stack([(x[t[:, 0]] + x[t[:, 1]] + x[t[:, 2]]) / 3.0,
                     (y[t[:, 0]] + y[t[:, 1]] + y[t[:, 2]]) / 3.0,
                     (z[t[:, 0]] + z[t[:, 1]] + z[t[:, 2]]) / 3.0], 1)


# ------------------------------------------------------------------ ellipsoid dome ------
def _polar_fan(v: int, rings: int):
    """Disk of `rings` rings around a centre of valence v, v triangular lattice sectors;
    returns 2-D points, CCW triangles and the rim in ring order."""
    ids = {}
    pts = [(0.0, 0.0)]

    def corner(f):
        a = 2 * np.pi * (f % v) / v
        return np.array([np.cos(a), np.sin(a)])

    def vid(f, i, j):
        f %= v
        if i == 0 and j == 0:
            return 0
        if i == 0:
            f, i, j = (f + 1) % v, j, 0
        key = (f, i, j)
        if key not in ids:
            p = i * corner(f) + j * corner(f + 1)
            ids[key] = len(pts)
            pts.append((p[0], p[1]))
        return ids[key]

    tris = []
    for f in range(v):
        for i in range(rings):
            for j in range(rings - i):
                tris.append((vid(f, i, j), vid(f, i + 1, j), vid(f, i, j + 1)))
                if i + j + 2 <= rings:
                    tris.append((vid(f, i + 1, j), vid(f, i + 1, j + 1), vid(f, i, j + 1)))
    rim = [vid(f, rings - k, k) for f in range(v) for k in range(rings)]
    return np.array(pts), [list(t) for t in tris], rim


def _flip_edges(P, T, rim, n_flips, rng):
    """Flip interior edges (convex quads only, valences kept in 4..8, rim untouched) to
    create mixed valence-5/7 extraordinary vertices in the coarse mesh."""
    rimset = set(rim)
    done = 0
    for _ in range(2000):
        if done >= n_flips:
            break
        edges = {}
        for ti, t in enumerate(T):
            for k in range(3):
                a, b = t[k], t[(k + 1) % 3]
                edges.setdefault((min(a, b), max(a, b)), []).append(ti)
        val = np.zeros(len(P), int)
        for t in T:
            for a in t:
                val[a] += 1
        cand = [e for e, ts in edges.items() if len(ts) == 2 and e[0] not in rimset and e[1] not in rimset]
        a, b = cand[rng.integers(len(cand))]
        t1, t2 = edges[(a, b)]
        c = [x for x in T[t1] if x not in (a, b)][0]
        d = [x for x in T[t2] if x not in (a, b)][0]
        if c in rimset or d in rimset:
            continue
        if min(val[a], val[b]) <= 5 or max(val[c], val[d]) >= 7:
            continue

        def cross(o, p, q):
            return (P[p][0] - P[o][0]) * (P[q][1] - P[o][1]) - (P[p][1] - P[o][1]) * (P[q][0] - P[o][0])
        # convexity of quad a-c-b-d: c and d on opposite sides of ab, a and b of cd
        if cross(c, d, a) * cross(c, d, b) >= 0 or cross(a, b, c) * cross(a, b, d) >= 0:
            continue
        # keep CCW orientation
        n1 = [a, d, c] if cross(a, d, c) > 0 else [a, c, d]
        n2 = [b, c, d] if cross(b, c, d) > 0 else [b, d, c]
        T[t1], T[t2] = n1, n2
        done += 1
    return T


def _loop_refine(P, T, rim):
    """One Loop subdivision step in 3-D (topology 1 -> 4): interior edge points
    3/8 (a + b) + 1/8 (c + d), interior vertices (1 - n b) v + b sum(nbrs) with Loop's
    b(n) = (5/8 - (3/8 + cos(2 pi / n) / 4)^2) / n; the rim follows the cubic B-spline curve
    rules (midpoints, 1/8 3/4 1/8).  New nodes: old vertices keep their ids, edge points
    follow in edge order; the four children of a triangle are contiguous."""
    T = np.asarray(T, dtype=np.int64)
    n = len(P)
    sides = np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]])
    opp = np.concatenate([T[:, 2], T[:, 0], T[:, 1]])
    key = np.minimum(sides[:, 0], sides[:, 1]) * n + np.maximum(sides[:, 0], sides[:, 1])
    ukey, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    a, b = ukey // n, ukey % n
    osum = np.zeros((len(ukey), 3))
    np.add.at(osum, inv, P[opp])
    bnd_edge = cnt == 1
    EP = np.where(bnd_edge[:, None], 0.5 * (P[a] + P[b]), 0.375 * (P[a] + P[b]) + 0.125 * osum)
    nsum = np.zeros((n, 3))
    val = np.zeros(n)
    np.add.at(nsum, a, P[b]); np.add.at(nsum, b, P[a])
    np.add.at(val, a, 1.0); np.add.at(val, b, 1.0)
    beta = (0.625 - (0.375 + 0.25 * np.cos(2 * np.pi / np.maximum(val, 1))) ** 2) / np.maximum(val, 1)
    VP = (1 - val * beta)[:, None] * P + beta[:, None] * nsum
    rim = np.asarray(rim)
    prev, nxt = np.roll(rim, 1), np.roll(rim, -1)
    VP[rim] = 0.75 * P[rim] + 0.125 * (P[prev] + P[nxt])
    eid = n + inv
    e01, e12, e20 = eid[: len(T)], eid[len(T): 2 * len(T)], eid[2 * len(T):]
    nt = np.stack([np.stack([T[:, 0], e01, e20], 1), np.stack([e01, T[:, 1], e12], 1),
                   np.stack([e20, e12, T[:, 2]], 1), np.stack([e01, e12, e20], 1)], 1).reshape(-1, 3)
    rk = np.minimum(rim, nxt) * n + np.maximum(rim, nxt)
    rmid = n + np.searchsorted(ukey, rk)
    nrim = np.stack([rim, rmid], 1).reshape(-1)
    return np.vstack([VP, EP]), nt, nrim


def ellipsoid_dome(levels: int = 6, a: float = 2.2, b: float = 1.0, z_cut: float = 0.5,
                   apex_valence: int = 5, coarse_rings: int = 3, n_flips: int = 4, seed: int = 3,
                   helix_deg=(-60.0, 60.0), stride: int = 18) -> Mesh:
    """Truncated prolate ellipsoid shell x^2/b^2 + y^2/b^2 + z^2/a^2 = 1, z >= z_cut.

    A coarse polar fan (apex valence `apex_valence`) gets `n_flips` random interior edge
    flips (default_rng(seed)), creating valence-5/7 vertices, and is refined `levels` times
    (1 -> 4) with Loop subdivision masks in 3-D, so the node positions are smooth,
    extraordinary vertices are isolated and every element touches at most one.  The refined
    net is projected onto the ellipsoid with a smooth polar-angle stretch that puts the rim
    on z = z_cut.  Rim nodes form one closed boundary ring with the reference's ghost
    layer (GlobalGhostInit rule).  Fibres follow the reference's LV construction
    (src/lv_geometry_unstructured.c:313-367) with the helix angle blended linearly from
    helix_deg[0] at the apex to helix_deg[1] at the rim.  levels = 6 gives 184 320 triangles.
    """
    rng = np.random.default_rng(seed)
    P2, T, rim = _polar_fan(apex_valence, coarse_rings)
    T = _flip_edges(P2, T, rim, n_flips, rng)
    th_rim = np.arccos(z_cut / a)

    def on_surface(th, ph):
        return np.stack([b * np.sin(th) * np.cos(ph), b * np.sin(th) * np.sin(ph), a * np.cos(th)], 1)

    # coarse control net on the surface (any placement: subdivision smooths it)
    ang = np.arctan2(P2[:, 1], P2[:, 0])
    P = on_surface(th_rim * np.hypot(P2[:, 0], P2[:, 1]) / coarse_rings, ang)
    for _ in range(levels):
        P, T, rim = _loop_refine(P, T, rim)
    # project onto the ellipsoid along (x/b, y/b, z/a), then stretch the polar angle
    # smoothly so that the rim lands on z = z_cut: th' = th (1 + (th_rim / th_b(ph) - 1) s),
    # s = (th / th_b(ph))^2, th_b the subdivided rim's polar angle interpolated linearly in
    # ph between rim nodes (exact on the rim, so no interior node is pushed past it).
    q = P / np.array([b, b, a])
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    th = np.arccos(np.clip(q[:, 2], -1.0, 1.0))
    ph = np.arctan2(q[:, 1], q[:, 0])
    rim_idx = np.asarray(rim)
    order = np.argsort(ph[rim_idx])
    thb = np.interp(ph, ph[rim_idx][order], th[rim_idx][order], period=2.0 * np.pi)
    sfac = (th / thb) ** 2
    th2 = th * (1.0 + (th_rim / thb - 1.0) * sfac)
    th2[rim] = th_rim
    X = on_surface(th2, ph)
    T = np.asarray(T, dtype=np.int32)
    n_v = len(X)
    rim = np.array(rim, dtype=np.int32)
    E = len(T)
    G = len(rim)
    lib = L.lib()
    nv = [np.zeros(E + 2 * G, np.int32) for _ in range(3)]
    for k in range(3):
        nv[k][:E] = T[:, k]
    rules = np.zeros(3 * G, np.int32)
    L.check("bf_topo_ghost_ring", lib.bf_topo_ghost_ring(
        n_v, E, 0, 1, L.ip(np.array([G], np.int32)), L.ip(rim), L.ip(nv[0]), L.ip(nv[1]),
        L.ip(nv[2]), L.ip(rules)))
    tri = np.stack(nv, 1).astype(np.int32)
    ire, irv, val = (np.zeros(E, np.int32) for _ in range(3))
    L.check("bf_topo_irrver", lib.bf_topo_irrver(n_v, E, 0, L.ip(_i32(T[:, 0])), L.ip(_i32(T[:, 1])),
                                                  L.ip(_i32(T[:, 2])), G, L.ip(rim), L.ip(ire),
                                                  L.ip(irv), L.ip(val)))
    n_nodes = n_v + G
    x0, y0, z0 = (np.zeros(n_nodes) for _ in range(3))
    x0[:n_v], y0[:n_v], z0[:n_v] = X[:, 0], X[:, 1], X[:, 2]
    m = Mesh(n_v=n_v, n_nodes=n_nodes, n_elmt=E, x0=x0, y0=y0, z0=z0, tri=tri,
             patch=np.zeros((E, stride), np.int32), ire=ire, irv=irv, val=val, dA0=np.zeros(E),
             n_fib=np.zeros((E, 3)), gamma_scale=np.ones(E), ghost_rules=rules.reshape(-1, 3),
             name=f"ellipsoid_L{levels}")
    m.apply_ghosts(m.x0, m.y0, m.z0)
    m.patch = build_patch(n_v, n_nodes, E, tri, ire, irv, val, stride)
    m.dA0 = area0(tri[:E], m.x0, m.y0, m.z0)
    # fibres (helix blended apex -> rim)
    c = centroids(m)
    e_n = np.stack([c[:, 0] / b**2, c[:, 1] / b**2, c[:, 2] / a**2], 1)
    e_n /= np.linalg.norm(e_n, axis=1, keepdims=True)
    down = np.array([0.0, 0.0, -1.0])
    e_l = down - (e_n @ down)[:, None] * e_n
    nl = np.linalg.norm(e_l, axis=1)
    fallback = np.array([1.0, 0.0, 0.0]) - (e_n @ np.array([1.0, 0, 0]))[:, None] * e_n
    e_l = np.where((nl > 1e-10)[:, None], e_l / np.maximum(nl, 1e-300)[:, None],
                   fallback / np.linalg.norm(fallback, axis=1, keepdims=True))
    e_c = np.cross(e_n, e_l)
    zstar = np.clip((a - c[:, 2]) / (a - z_cut), 0.0, 1.0)
    alpha = np.deg2rad((1 - zstar) * helix_deg[0] + zstar * helix_deg[1])
    m.n_fib = np.cos(alpha)[:, None] * e_c + np.sin(alpha)[:, None] * e_l
    m.extra.update(levels=levels, a=a, b=b, z_cut=z_cut, rim=rim)
    return m
