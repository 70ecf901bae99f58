"""The element tiling of the single-pass staged kernel (biofem-ia_amd/csrc/bf_tiles.cpp),
checked on the CPU: every index the kernels follow stays inside its tile and buffers, every
(element, slot) contribution is listed exactly once, under the local node its slot code
names, and replaying the kernels' summation order (each local node's contributions in list
order, complete nodes written by their tile, partial nodes summed over their consecutive
tile partials in tile order) reproduces the direct node sums of arbitrary element
contributions."""
import ctypes as C
import os

import numpy as np
import pytest

from biofem_amd import configs
from biofem_amd.mesh import build_patch, flat_membrane

HERE = os.path.dirname(os.path.abspath(__file__))
SENTINEL = 1000000
IDLE = 0xFFFF


def node_cap(T):
    return 2 * T + T // 2 + 64


@pytest.fixture(scope="module")
def shim():
    path = os.path.join(HERE, "libelem_shim.so")
    if not os.path.exists(path):
        pytest.skip("tests/libelem_shim.so not built (run __graft_entry__.build())")
    return C.CDLL(path)


def device_patch(m):
    """The patch table the library tiles: slot-major [NS][E], NS the widest patch, empty or
    unused slots parked on node 0, and -1 in the slots of an LV apex-cap element other than
    its three corners (they add no force)."""
    nen = m.nen()
    NS = int(nen.max())
    E = m.n_elmt
    P = np.zeros((NS, E), np.int32)
    for i in range(min(NS, m.patch.shape[1])):
        col = m.patch[:E, i].astype(np.int64)
        ok = (i < nen) & (col != SENTINEL) & (col >= 0) & (col < m.n_nodes)
        P[i] = np.where(ok, col, 0)
    for e in range(m.n_elmt_base if m.n_elmt_base > 0 else E, E):
        nslot = int(m.val[e]) + 6
        corners = [next(i for i in range(nslot) if m.patch[e, i] == c) for c in m.tri[e]]
        for i in range(nen[e]):
            if i not in corners:
                P[i, e] = -1
    return P, nen.astype(np.int32), NS


def tiles(shim, m, T):
    P, nen, NS = device_patch(m)
    E = m.n_elmt
    ip = C.POINTER(C.c_int)
    dp = C.POINTER(C.c_double)
    args = [E, m.n_v, m.n_nodes, NS, P.ctypes.data_as(ip), nen.ctypes.data_as(ip),
            m.x0.ctypes.data_as(dp), m.y0.ctypes.data_as(dp), m.z0.ctypes.data_as(dp), T]
    sizes = np.zeros(4, np.int32)
    assert shim.shim_tiles(*args, sizes.ctypes.data_as(ip), *([None] * 9)) == 0
    nt, nl, npart, nc = (int(v) for v in sizes)
    out = dict(elem=np.zeros(nt * T, np.int32), lslot=np.zeros(nt * NS * T, np.uint16),
               node_ptr=np.zeros(nt + 1, np.int32), node=np.zeros(max(nl, 1), np.int32),
               out=np.zeros(max(nl, 1), np.int32), cptr=np.zeros(nl + 1, np.int32),
               cent=np.zeros(max(nc, 1), np.uint16), pptr=np.zeros(m.n_v + 1, np.int32),
               complete=np.zeros(max(m.n_v, 1), np.uint8))
    assert shim.shim_tiles(*args, sizes.ctypes.data_as(ip),
                           *[out[k].ctypes.data_as(C.c_void_p) for k in
                             ("elem", "lslot", "node_ptr", "node", "out", "cptr", "cent", "pptr",
                              "complete")]) == 0
    out.update(n_tiles=nt, NS=NS, n_partials=npart, n_cent=nc, P=P, nen=nen)
    return out


def meshes():
    ms = {"flat 64x48": flat_membrane(64, 48, 1.0, 0.75, jitter=1e-3, seed=0)}
    ms["ellipsoid L4"] = configs.ellipsoid_shell(levels=4)[0]
    cap = flat_membrane(12, 9, 1.0, 0.75, jitter=1e-3, seed=0)
    base = cap.n_elmt - 20
    cap.n_elmt_base = base
    cap.patch = build_patch(cap.n_v, cap.n_nodes, cap.n_elmt, cap.tri, cap.ire, cap.irv,
                            cap.val, 16, n_elmt_base=base)
    ms["lv cap"] = cap
    return ms


@pytest.mark.parametrize("T", [64, 128])
@pytest.mark.parametrize("name", ["flat 64x48", "ellipsoid L4", "lv cap"])
def test_tiling_invariants_and_summation(shim, name, T):
    m = meshes()[name]
    t = tiles(shim, m, T)
    E, NS, nt = m.n_elmt, t["NS"], t["n_tiles"]
    P, nen = t["P"], t["nen"]
    elem = t["elem"].reshape(nt, T)
    lslot = t["lslot"].reshape(nt, NS, T)
    nptr, node, out, cptr, cent = t["node_ptr"], t["node"], t["out"], t["cptr"], t["cent"]
    # every element in exactly one lane
    live = elem[elem >= 0]
    assert np.array_equal(np.sort(live), np.arange(E))
    assert cptr[0] == 0 and cptr[-1] == t["n_cent"] and (np.diff(cptr) >= 0).all()
    rng = np.random.default_rng(7)
    contrib = rng.standard_normal((E, NS, 3))           # element e's force on slot a
    partial = np.full((max(t["n_partials"], 1), 3), np.nan)
    fint = np.full((m.n_nodes, 3), np.nan)
    for k in range(nt):
        j0, L = nptr[k], nptr[k + 1] - nptr[k]
        assert 0 < L <= node_cap(T)
        loc_nodes = node[j0:j0 + L]
        assert len(set(loc_nodes.tolist())) == L
        used = lslot[k] != IDLE
        for a in range(NS):
            for lane in range(T):
                e = elem[k, lane]
                assert used[a, lane] == (e >= 0 and a < nen[e] and P[a, e] >= 0)
        assert (lslot[k][used] < L).all()
        for a in range(NS):
            assert np.array_equal(loc_nodes[lslot[k, a][used[a]]], P[a, elem[k, used[a]]])
        # the node lists hold every used (lane, slot) once, under its slot's local node,
        # in lane order, then slot order
        seen = set()
        for li in range(L):
            ent = cent[cptr[j0 + li]:cptr[j0 + li + 1]]
            lanes, slots = ent & 0xFF, ent >> 8
            assert (lanes < T).all() and (slots < NS).all()
            key = lanes.astype(np.int64) * 64 + slots
            assert (np.diff(key) > 0).all()
            for lane, a in zip(lanes, slots):
                assert lslot[k, a, lane] == li
                seen.add((int(lane), int(a)))
            acc = np.zeros(3)
            for lane, a in zip(lanes, slots):                          # list order
                acc = acc + contrib[elem[k, lane], a]
            n, o = loc_nodes[li], out[j0 + li]
            if n >= m.n_v:
                assert o == -2
            elif o == -1:
                assert t["complete"][n] == 1
                fint[n] = acc
            else:
                assert t["pptr"][n] <= o < t["pptr"][n + 1] and t["complete"][n] == 0
                partial[o] = acc
        assert seen == {(int(lane), int(a)) for a, lane in zip(*np.nonzero(used))}
    # node pass: a partial node's consecutive partials, in tile order
    pptr = t["pptr"]
    for n in range(m.n_v):
        if pptr[n + 1] > pptr[n]:
            f = np.zeros(3)
            for j in range(pptr[n], pptr[n + 1]):
                f = f + partial[j]
            fint[n] = f
    direct = np.zeros((m.n_nodes, 3))
    valid = (np.arange(NS)[:, None] < nen[None, :]) & (P >= 0)
    for a in range(NS):
        ok = valid[a]
        np.add.at(direct, P[a, ok], contrib[ok, a])
    real = np.zeros(m.n_nodes, bool)
    real[:m.n_v] = True
    touched = np.zeros(m.n_nodes, bool)
    touched[P[valid]] = True
    chk = real & touched
    np.testing.assert_allclose(fint[chk], direct[chk], rtol=1e-12, atol=1e-12)
    print(f"  {name} T={T}: {nt} tiles, {len(node)} local nodes, {t['n_partials']} partials")
