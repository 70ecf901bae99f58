"""The C host layer (include/biofem_fe.h): bf_FInternalAct on a bf_fe struct behaves like
the reference's FInternalAct(FE *): ADD into Fint, ghost rows untouched."""
import ctypes as C

import numpy as np
import pytest

from biofem_amd import _lib as L
from biofem_amd import configs
from oracle import bf_oracle as O

pytestmark = pytest.mark.gpu
_ip, _dp = C.POINTER(C.c_int), C.POINTER(C.c_double)


class Ibm(C.Structure):
    _fields_ = [("n_v", C.c_int), ("n_ghosts", C.c_int), ("n_elmt", C.c_int), ("n_elmt_base", C.c_int),
                ("x_bp", _dp), ("y_bp", _dp), ("z_bp", _dp), ("x_bp0", _dp), ("y_bp0", _dp), ("z_bp0", _dp),
                ("nv1", _ip), ("nv2", _ip), ("nv3", _ip), ("patch", _ip), ("patch_stride", C.c_int),
                ("ire", _ip), ("val", _ip), ("dA0", _dp), ("n_fib", _dp), ("gamma_scale", _dp),
                ("activation_delay", _dp), ("n_ghost_rules", C.c_int), ("ghost_rules", _ip)]


class Fe(C.Structure):
    _fields_ = [("ibm", C.POINTER(Ibm)), ("act", L.BfParams), ("Fint", _dp), ("precision", C.c_int),
                ("device", C.c_int), ("ctx", C.c_void_p)]


def test_finternalact_adds_like_the_reference():
    lib = L.lib()
    lib.bf_FInternalAct.argtypes = [C.POINTER(Fe)]
    lib.bf_FEDestroy.argtypes = [C.POINTER(Fe)]
    lib.bf_GlobalGhost.argtypes = [C.POINTER(Fe)]
    m, p, _ = configs.membrane_small()
    x, y, z = configs.smooth_displacement(m, 0.01)
    keep = []

    def d(a):
        a = np.ascontiguousarray(a, np.float64)
        keep.append(a)
        return a.ctypes.data_as(_dp)

    def i(a):
        a = np.ascontiguousarray(a, np.int32)
        keep.append(a)
        return a.ctypes.data_as(_ip)

    xs, ys, zs = x.copy(), y.copy(), z.copy()
    xs[m.n_v:] = ys[m.n_v:] = zs[m.n_v:] = 0.0   # ghosts stale: bf_GlobalGhost refreshes them
    ibm = Ibm(n_v=m.n_v, n_ghosts=m.n_nodes - m.n_v, n_elmt=m.n_elmt, n_elmt_base=0,
              x_bp=xs.ctypes.data_as(_dp), y_bp=ys.ctypes.data_as(_dp), z_bp=zs.ctypes.data_as(_dp),
              x_bp0=d(m.x0), y_bp0=d(m.y0), z_bp0=d(m.z0), nv1=i(m.tri[:, 0]), nv2=i(m.tri[:, 1]),
              nv3=i(m.tri[:, 2]), patch=i(m.patch), patch_stride=16, ire=i(m.ire), val=i(m.val),
              dA0=d(m.dA0), n_fib=d(m.n_fib.reshape(-1)), gamma_scale=d(m.gamma_scale),
              activation_delay=None, n_ghost_rules=len(m.ghost_rules), ghost_rules=i(m.ghost_rules.reshape(-1)))
    fint = np.full(3 * m.n_nodes, 0.25)
    fe = Fe(ibm=C.pointer(ibm), act=p.to_c(), Fint=fint.ctypes.data_as(_dp), precision=L.BF_PREC_FP64,
            device=0, ctx=None)
    L.check("ghost", lib.bf_GlobalGhost(C.byref(fe)))
    assert np.array_equal(xs, x) and np.array_equal(zs, z)
    L.check("fint", lib.bf_FInternalAct(C.byref(fe)))
    ref, _, _ = O.finternal_act(m, p, x, y, z, fint=np.full(3 * m.n_nodes, 0.25), accumulate=True)
    assert np.array_equal(fint, ref)
    L.check("destroy", lib.bf_FEDestroy(C.byref(fe)))
