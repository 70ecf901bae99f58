"""ctypes wrapper of the CPU oracle (oracle/bf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point as the checker and
by bench.py's cpu_baseline leg.  The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class Mesh(C.Structure):
    _fields_ = [
        ("n_v", C.c_int), ("n_nodes", C.c_int), ("n_elmt", C.c_int), ("n_elmt_base", C.c_int),
        ("patch_stride", C.c_int), ("patch", _ip), ("ire", _ip), ("val", _ip),
        ("nv1", _ip), ("nv2", _ip), ("nv3", _ip),
        ("x0", _dp), ("y0", _dp), ("z0", _dp), ("dA0", _dp), ("n_fib", _dp),
        ("gamma_scale", _dp), ("activation_delay", _dp),
    ]


class Params(C.Structure):
    _fields_ = [
        ("n_qp", C.c_int), ("c33_iters", C.c_int), ("cart_fib_act", C.c_int),
        ("manufactured", C.c_int), ("h0", C.c_double), ("mu", C.c_double), ("K", C.c_double),
        ("gamma", C.c_double), ("ramp_type", C.c_int), ("gamma_T", C.c_double),
        ("tau_up", C.c_double), ("tau_down", C.c_double), ("apd", C.c_double),
        ("pulse_t0", C.c_double), ("pulse_sigma", C.c_double),
        ("manufactured_gamma0", C.c_double), ("manufactured_T", C.c_double),
        ("ti", C.c_int), ("tisteps", C.c_int), ("dt", C.c_double),
    ]


class Status(C.Structure):
    _fields_ = [("status", C.c_int), ("elem", C.c_int), ("n_warn", C.c_int)]


class Geom(C.Structure):
    _fields_ = [("nen", C.c_int), ("v", C.c_int), ("is_irregular", C.c_int),
                ("vec", C.c_double * 24),
                ("INa0", C.c_double * 64), ("INa1", C.c_double * 64), ("INab0", C.c_double * 64),
                ("INab1", C.c_double * 64), ("INab2", C.c_double * 64)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}


def lib(omp: bool = False) -> C.CDLL:
    name = "libbf_oracle_omp.so" if omp else "libbf_oracle.so"
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.bfo_finternal_act.argtypes = [C.POINTER(Mesh), C.POINTER(Params), _dp, _dp, _dp, _dp,
                                        C.c_int, _dp, C.POINTER(Status)]
        L.bfo_finternal_act_omp.argtypes = L.bfo_finternal_act.argtypes + [C.c_int]
        L.bfo_gauss.argtypes = [C.c_int, _dp, _dp]
        L.bfo_gamma_of_time.argtypes = [C.POINTER(Params), C.c_double]
        L.bfo_gamma_of_time.restype = C.c_double
        L.bfo_irregular_tables.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.bfo_geom_subdiv.argtypes = [C.POINTER(Mesh), C.c_int, _dp, _dp, _dp, C.POINTER(Geom)]
        L.bfo_area_normal.argtypes = [C.c_int, _ip, _ip, _ip, _dp, _dp, _dp, _dp]
        L.bfo_irrver.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, C.c_int, _ip, _ip, _ip, _ip]
        L.bfo_patch.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip,
                                C.c_int, _ip]
        L.bfo_global_ghost_init.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip,
                                            _ip, _dp, _dp, _dp, _dp, _dp, _dp]
        L.bfo_global_ghost.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _ip, _ip, _ip,
                                       _dp, _dp, _dp]
        _libs[name] = L
    return _libs[name]


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


class OracleMesh:
    """Keeps the numpy arrays alive behind a bfo_mesh."""

    def __init__(self, m):
        keep = []

        def ia(a):
            arr, p = _i(a)
            keep.append(arr)
            return p

        def da(a):
            arr, p = _d(a)
            keep.append(arr)
            return p

        s = Mesh()
        s.n_v, s.n_nodes, s.n_elmt = m.n_v, m.n_nodes, m.n_elmt
        s.n_elmt_base = m.n_elmt_base
        s.patch_stride = m.patch.shape[1]
        s.patch = ia(m.patch)
        s.ire, s.val = ia(m.ire), ia(m.val)
        t = m.tri[: m.n_elmt]
        s.nv1, s.nv2, s.nv3 = ia(t[:, 0]), ia(t[:, 1]), ia(t[:, 2])
        s.x0, s.y0, s.z0 = da(m.x0), da(m.y0), da(m.z0)
        s.dA0 = da(m.dA0)
        s.n_fib = da(np.asarray(m.n_fib).reshape(-1))
        s.gamma_scale = da(m.gamma_scale)
        s.activation_delay = da(m.activation_delay) if m.activation_delay is not None else None
        self.s = s
        self.keep = keep


def params(p) -> Params:
    q = Params()
    for name, _ in Params._fields_:
        setattr(q, name, getattr(p, name))
    return q


def finternal_act(m, p, x, y, z, fint=None, accumulate=False, omp_threads=0, want_c33=False):
    """Oracle FInternalAct; returns (fint, status, c33)."""
    om = OracleMesh(m)
    pp = params(p)
    x, xp = _d(x)
    y, yp = _d(y)
    z, zp = _d(z)
    if fint is None:
        fint = np.zeros(3 * m.n_nodes)
        accumulate = False
    fint = np.ascontiguousarray(fint, dtype=np.float64)
    c33 = np.zeros(m.n_elmt * p.n_qp) if want_c33 else None
    st = Status()
    fp = fint.ctypes.data_as(_dp)
    cp = c33.ctypes.data_as(_dp) if c33 is not None else None
    if omp_threads:
        lib(True).bfo_finternal_act_omp(C.byref(om.s), C.byref(pp), xp, yp, zp, fp,
                                        int(accumulate), cp, C.byref(st), int(omp_threads))
    else:
        lib().bfo_finternal_act(C.byref(om.s), C.byref(pp), xp, yp, zp, fp, int(accumulate), cp,
                                C.byref(st))
    return fint, (st.status, st.elem, st.n_warn), (c33.reshape(m.n_elmt, p.n_qp) if c33 is not None else None)


def irregular_tables(v: int) -> np.ndarray:
    out = np.zeros((5, 64))
    ptrs = [out[i].ctypes.data_as(_dp) for i in range(5)]
    lib().bfo_irregular_tables(v, *ptrs)
    return out[:, : v + 6].copy()


def gauss(n: int):
    th, w = np.zeros(5), np.zeros(5)
    rc = lib().bfo_gauss(n, th.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    return rc, th[:n], w[:n]


def gamma_of_time(p, delay=0.0) -> float:
    return lib().bfo_gamma_of_time(C.byref(params(p)), float(delay))


def geom(m, ec, x, y, z):
    """bfo_geom_subdiv of one element: dict of the SubdivGeomQP vectors."""
    om = OracleMesh(m)
    g = Geom()
    x, xp = _d(x)
    y, yp = _d(y)
    z, zp = _d(z)
    rc = lib().bfo_geom_subdiv(C.byref(om.s), ec, xp, yp, zp, C.byref(g))
    v = np.array(g.vec[:]).reshape(8, 3)
    names = ["ndx21", "ndx31", "nn", "gc1", "gc2", "Aaa", "Abb", "Aab"]
    out = {k: v[i] for i, k in enumerate(names)}
    out.update(rc=rc, nen=g.nen, v=g.v, is_irregular=g.is_irregular,
               IN=np.array([g.INa0[: g.nen], g.INa1[: g.nen], g.INab0[: g.nen], g.INab1[: g.nen],
                            g.INab2[: g.nen]]))
    return out


def irrver(n_v, n_elmt, tri, bnodes, n_elmt_base=0):
    t, _ = _i(tri)
    a, ap = _i(t[:, 0])
    b, bp = _i(t[:, 1])
    c, cp = _i(t[:, 2])
    bn, bnp = _i(bnodes if len(bnodes) else np.zeros(1, np.int32))
    ire = np.zeros(n_elmt, np.int32)
    irv = np.zeros(n_elmt, np.int32)
    val = np.zeros(n_elmt, np.int32)
    lib().bfo_irrver(n_v, n_elmt, n_elmt_base, ap, bp, cp, len(bnodes), bnp,
                     ire.ctypes.data_as(_ip), irv.ctypes.data_as(_ip), val.ctypes.data_as(_ip))
    return ire, irv, val


def patch(n_v, n_elmt, tri, ire, irv, val, stride=16, n_elmt_base=0):
    t, _ = _i(tri)
    a, ap = _i(t[:, 0])
    b, bp = _i(t[:, 1])
    c, cp = _i(t[:, 2])
    ire, irep = _i(ire)
    irv, irvp = _i(irv)
    val, valp = _i(val)
    out = np.zeros((n_elmt, stride), np.int32)
    lib().bfo_patch(n_v, n_elmt, n_elmt_base, len(t), ap, bp, cp, irep, irvp, valp, stride,
                    out.ctypes.data_as(_ip))
    return out


def area_normal(tri, x, y, z):
    t, _ = _i(tri)
    a, ap = _i(t[:, 0])
    b, bp = _i(t[:, 1])
    c, cp = _i(t[:, 2])
    x, xp = _d(x)
    y, yp = _d(y)
    z, zp = _d(z)
    dA = np.zeros(len(t))
    lib().bfo_area_normal(len(t), ap, bp, cp, xp, yp, zp, dA.ctypes.data_as(_dp))
    return dA


class Cached:
    """geom0 built once, then repeated FInternalAct evaluations (CPU baseline timing)."""

    def __init__(self, m, p, omp: bool = False):
        self.L = lib(omp)
        self.L.bfo_geom0_create.restype = C.c_void_p
        self.L.bfo_geom0_create.argtypes = [C.POINTER(Mesh), C.POINTER(Status)]
        self.L.bfo_geom0_free.argtypes = [C.c_void_p]
        self.L.bfo_finternal_act_cached.argtypes = [C.POINTER(Mesh), C.POINTER(Params), C.c_void_p,
                                                    _dp, _dp, _dp, _dp, C.c_int, _dp,
                                                    C.POINTER(Status), C.c_int]
        self.m, self.om, self.pp = m, OracleMesh(m), params(p)
        st = Status()
        self.g0 = self.L.bfo_geom0_create(C.byref(self.om.s), C.byref(st))
        self.status = (st.status, st.elem, st.n_warn)

    def eval(self, x, y, z, fint, nthreads=1):
        st = Status()
        self.L.bfo_finternal_act_cached(C.byref(self.om.s), C.byref(self.pp), self.g0,
                                        x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                                        z.ctypes.data_as(_dp), fint.ctypes.data_as(_dp), 0, None,
                                        C.byref(st), int(nthreads))
        return (st.status, st.elem, st.n_warn)

    def close(self):
        if self.g0:
            self.L.bfo_geom0_free(self.g0)
            self.g0 = None


def element_force(m, p, ec, x, y, z):
    """bfo_element_force of one element: (Fb[3*nen], S[n_qp,3,3], C33[n_qp], n_warn)."""
    L = lib()
    L.bfo_element_force.argtypes = [C.POINTER(Mesh), C.POINTER(Params), C.POINTER(Geom),
                                    C.POINTER(Geom), C.c_int, _dp, _dp, _dp, C.POINTER(C.c_int)]
    om = OracleMesh(m)
    g0, gc = Geom(), Geom()
    x0, x0p = _d(m.x0)
    y0, y0p = _d(m.y0)
    z0, z0p = _d(m.z0)
    L.bfo_geom_subdiv(C.byref(om.s), ec, x0p, y0p, z0p, C.byref(g0))
    x, xp = _d(x)
    y, yp = _d(y)
    z, zp = _d(z)
    L.bfo_geom_subdiv(C.byref(om.s), ec, xp, yp, zp, C.byref(gc))
    Fb = np.zeros(3 * 64)
    S = np.zeros(9 * 5)
    c33 = np.zeros(5)
    w = C.c_int(0)
    L.bfo_element_force(C.byref(om.s), C.byref(params(p)), C.byref(g0), C.byref(gc), ec,
                        Fb.ctypes.data_as(_dp), S.ctypes.data_as(_dp), c33.ctypes.data_as(_dp),
                        C.byref(w))
    nen = 12 if m.ire[ec] == 0 else int(m.val[ec]) + 6
    return Fb[: 3 * nen], S[: 9 * p.n_qp].reshape(p.n_qp, 3, 3), c33[: p.n_qp], w.value


def explicit_steps(m, p, x, y, z, mass, fixed, dt, damp, n_steps, v=None, nthreads=4):
    """Host restatement of the library's explicit damped time step (bf_step), for tests:
        GlobalGhost (x[n_v+g] = (x[b] + x[a]) - x[c]), f = finternal_act(x),
        v <- damp * v - (dt / m_n) * f,  x <- x + dt * v   on free real nodes,
    every product and sum rounded separately (numpy element-wise, no contraction).  This
    step is the library's substitute for the reference's implicit Newmark/SNES step
    (src/main.c:461-502), which needs PETSc: its parity against the reference is unpinned;
    this function pins the GPU trajectory against the oracle's force."""
    x, y, z = (np.array(a, dtype=np.float64, copy=True) for a in (x, y, z))
    n, nv = m.n_nodes, m.n_v
    v = np.zeros((n, 3)) if v is None else np.array(v, dtype=np.float64).reshape(n, 3).copy()
    free = np.zeros(n, bool)
    free[:nv] = ~np.asarray(fixed[:nv], bool)
    w = np.zeros(n)
    w[free] = dt / np.asarray(mass)[free]
    cached = Cached(m, p, omp=True)   # -O3 build, bit-identical to the -O2 one
    f = np.zeros(3 * n)
    try:
        for _ in range(n_steps):
            m.apply_ghosts(x, y, z)
            cached.eval(x, y, z, f, nthreads=nthreads)
            fr = f.reshape(n, 3)
            for c, a in enumerate((x, y, z)):
                vc = damp * v[free, c] - w[free] * fr[free, c]
                v[free, c] = vc
                a[free] = a[free] + dt * vc
    finally:
        cached.close()
    return x, y, z, v
