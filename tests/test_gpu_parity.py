"""GPU parity of the HIP force path against the CPU oracle (through the C ABI).

Tolerances: plain FP64 path <= 1e-12 relative (2-norm over the force vector) against the
oracle's FP64 restatement of FInternalAct; FP32-mixed <= 1e-5.
"""
import numpy as np
import pytest

from biofem_amd import _lib as L
from biofem_amd import configs
from biofem_amd.fem import ActStrainFE
from oracle import bf_oracle as O

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_MIXED = 1e-5


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def small():
    m, p, xyz = configs.membrane_small()
    fe = ActStrainFE(m, p)
    yield m, p, xyz, fe
    fe.close()


def test_membrane_small_fp64(small):
    m, p, (x, y, z), fe = small
    print("case membrane 32x24 fp64", flush=True)
    ref, st, c33 = O.finternal_act(m, p, x, y, z, want_c33=True)
    got = fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64)
    assert rel(got, ref) <= TOL_FP64
    assert np.all(got[3 * m.n_v:] == 0.0)
    np.testing.assert_allclose(fe.c33(), c33, rtol=0, atol=1e-12)
    assert fe.last_error() == (-1, 0, 0)


def test_membrane_small_displaced(small):
    m, p, _, fe = small
    x, y, z = configs.smooth_displacement(m)
    ref, _, _ = O.finternal_act(m, p, x, y, z)
    got = fe.finternal_act(x, y, z)
    assert rel(got, ref) <= TOL_FP64
    gm = fe.finternal_act(x, y, z, precision=L.BF_PREC_MIXED)
    assert rel(gm, ref) <= TOL_MIXED


def test_membrane_small_mixed(small):
    m, p, (x, y, z), fe = small
    ref, _, _ = O.finternal_act(m, p, x, y, z)
    got = fe.finternal_act(x, y, z, precision=L.BF_PREC_MIXED)
    assert rel(got, ref) <= TOL_MIXED


def test_deterministic_and_accumulate(small):
    m, p, (x, y, z), fe = small
    a = fe.finternal_act(x, y, z)
    b = fe.finternal_act(x, y, z)
    assert np.array_equal(a, b)
    base = np.arange(3 * m.n_nodes, dtype=np.float64) * 1e-3
    acc = fe.finternal_act(x, y, z, fint=base.copy(), accumulate=True)
    ref, _, _ = O.finternal_act(m, p, x, y, z, fint=base.copy(), accumulate=True)
    assert rel(acc, ref) <= TOL_FP64
    assert np.array_equal(acc[3 * m.n_v:], base[3 * m.n_v:])


def test_ghost_update(small):
    m, p, _, fe = small
    x, y, z = configs.smooth_displacement(m)
    gx, gy, gz = x.copy(), y.copy(), z.copy()
    gx[m.n_v:] = gy[m.n_v:] = gz[m.n_v:] = -7.0
    fe.ghost_update(gx, gy, gz)
    for a, b in ((gx, x), (gy, y), (gz, z)):
        assert np.array_equal(a, b)
