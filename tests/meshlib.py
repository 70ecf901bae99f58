"""Test-only mesh helpers: an independent Loop subdivision step and a polar fan mesh with
one extraordinary vertex.  Used to pin the irregular-patch evaluation of the oracle against
the definition of Loop subdivision (vertex mask (1 - n b) v + b sum(neighbours) with Loop's
b(n) = (5/8 - (3/8 + cos(2 pi / n) / 4)^2) / n, edge mask 3/8, 3/8, 1/8, 1/8)."""
import numpy as np


def fan_mesh(v: int, rings: int):
    """Disk of `rings` rings around a centre vertex of valence v; every other interior vertex
    has valence 6.  Returns (points [n,3], triangles [m,3], boundary node list in ring order)."""
    ids = {}
    pts = [(0.0, 0.0, 0.0)]

    def corner(f):
        a = 2 * np.pi * (f % v) / v
        return np.array([np.cos(a), np.sin(a)])

    def vid(f, i, j):
        # lattice point i c_f + j c_{f+1} of face f; (0, j) of face f is (j, 0) of face f+1
        f %= v
        if i == 0 and j == 0:
            return 0
        if i == 0:
            f, i, j = (f + 1) % v, j, 0
        key = (f, i, j)
        if key not in ids:
            p = i * corner(f) + j * corner(f + 1)
            ids[key] = len(pts)
            pts.append((p[0], p[1], 0.0))
        return ids[key]

    tris = []
    for f in range(v):
        for i in range(rings):
            for j in range(rings - i):
                tris.append((vid(f, i, j), vid(f, i + 1, j), vid(f, i, j + 1)))
                if i + j + 2 <= rings:
                    tris.append((vid(f, i + 1, j), vid(f, i + 1, j + 1), vid(f, i, j + 1)))
    bnd = []
    for f in range(v):
        for k in range(rings):
            bnd.append(vid(f, rings - k, k))
    return np.array(pts), np.array(tris, dtype=np.int32), np.array(bnd, dtype=np.int32)


def loop_subdivide(P: np.ndarray, T: np.ndarray):
    """One Loop subdivision step of an interior region (boundary rules not needed: only
    interior refined nodes are used by the tests).  Returns (points, triangles, edge_id)
    where edge_id[(a, b)] is the index of the edge point of edge (a, b)."""
    n = len(P)
    nbr = [set() for _ in range(n)]
    opp = {}
    for t in T:
        for k in range(3):
            a, b, c = t[k], t[(k + 1) % 3], t[(k + 2) % 3]
            nbr[a].add(b)
            nbr[b].add(a)
            opp.setdefault((min(a, b), max(a, b)), []).append(c)
    newP = np.zeros((n, 3))
    for i in range(n):
        k = len(nbr[i])
        b = (1.0 / k) * (0.625 - (0.375 + 0.25 * np.cos(2 * np.pi / k)) ** 2)
        newP[i] = (1 - k * b) * P[i] + b * sum(P[j] for j in nbr[i])
    eid = {}
    extra = []
    for (a, b), cs in opp.items():
        eid[(a, b)] = eid[(b, a)] = n + len(extra)
        if len(cs) == 2:
            extra.append(0.375 * (P[a] + P[b]) + 0.125 * (P[cs[0]] + P[cs[1]]))
        else:
            extra.append(0.5 * (P[a] + P[b]))
    allP = np.vstack([newP, np.array(extra)])
    nt = []
    for a, b, c in T:
        ab, bc, ca = eid[(a, b)], eid[(b, c)], eid[(c, a)]
        nt += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
    return allP, np.array(nt, dtype=np.int32), eid
