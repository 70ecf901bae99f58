"""Time stepping on resident state (the explicit damped step that replaces the reference's
PETSc-solved implicit step): 100 steps of the 32 x 24 membrane in the reference-order FP64
path reproduce the oracle-driven host restatement bit for bit; the staged paths stay
within their budgets along the trajectory."""
import ctypes as C

import numpy as np
import pytest

from biofem_amd import _lib as L
from biofem_amd import configs
from biofem_amd.batch import Batch, upload_coords
from biofem_amd.fem import ActStrainFE
from biofem_amd.mesh import lumped_mass
from biofem_amd.serve import stable_dt
from oracle import bf_oracle as O

pytestmark = pytest.mark.gpu


def _setup(m, p, steps_dt_scale=1.0):
    mass = lumped_mass(m, 1.0, p.h0)
    fixed = np.zeros(m.n_nodes, np.uint8)
    fixed[: m.n_v][np.isclose(m.x0[: m.n_v], 0.0)] = 1   # clamped edge x = 0
    dt = stable_dt(m, p) * steps_dt_scale
    return mass, fixed, dt, 0.995


def _run(m, p, mass, fixed, dt, damp, n, prec):
    fe = ActStrainFE(m, p)
    try:
        upload_coords(fe, m.x0, m.y0, m.z0)
        L.check("state", fe.lib.bf_state_init(fe.ctx, L.dp(mass), fixed.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                              None, dt, damp))
        L.check("step", fe.lib.bf_step(fe.ctx, prec, n))
        x, y, z = (np.zeros(m.n_nodes) for _ in range(3))
        v = np.zeros(3 * m.n_nodes)
        L.check("download", fe.lib.bf_state_download(fe.ctx, L.dp(x), L.dp(y), L.dp(z), L.dp(v)))
        return x, y, z, v
    finally:
        fe.close()


def test_hundred_steps_match_oracle():
    m, p, _ = configs.membrane_small()
    mass, fixed, dt, damp = _setup(m, p)
    print("case membrane 32x24, 100 explicit steps, dt", dt, flush=True)
    xr, yr, zr, vr = O.explicit_steps(m, p, m.x0, m.y0, m.z0, mass, fixed, dt, damp, 100)
    x, y, z, v = _run(m, p, mass, fixed, dt, damp, 100, L.BF_PREC_FP64)
    n = m.n_v
    assert np.array_equal(x[:n], xr[:n]) and np.array_equal(y[:n], yr[:n]) and np.array_equal(z[:n], zr[:n])
    assert np.array_equal(v.reshape(-1, 3)[:n], vr[:n])
    disp = np.linalg.norm(np.stack([xr - m.x0, yr - m.y0, zr - m.z0])[:, :n])
    assert disp > 0.0
    for prec, tol in ((L.BF_PREC_FP64_FAST, 1e-10), (L.BF_PREC_MIXED, 1e-5)):
        xs, ys, zs, _ = _run(m, p, mass, fixed, dt, damp, 100, prec)
        err = np.linalg.norm(np.stack([xs - xr, ys - yr, zs - zr])[:, :n]) / disp
        print(f"  prec {prec}: displacement rel err {err:.3e}", flush=True)
        assert err <= tol


def test_batched_steps_equal_single():
    cases = configs.batch_membranes(n=4, seed=3, nx_lo=8, nx_hi=24)
    print("case batched steps", flush=True)
    fes, ref = [], []
    try:
        for m, p, _ in cases:
            mass, fixed, dt, damp = _setup(m, p)
            ref.append(_run(m, p, mass, fixed, dt, damp, 30, L.BF_PREC_MIXED))
            fe = ActStrainFE(m, p)
            upload_coords(fe, m.x0, m.y0, m.z0)
            L.check("state", fe.lib.bf_state_init(fe.ctx, L.dp(mass),
                                                  fixed.ctypes.data_as(C.POINTER(C.c_ubyte)), None, dt, damp))
            fes.append(fe)
        b = Batch(fes)
        for k in (7, 13, 10):      # 30 steps in uneven rounds, graph replays
            L.check("bstep", b.lib.bf_batch_step(b.h, L.BF_PREC_MIXED, k, 1))
        b.sync()
        for fe, (m, _, _), r in zip(fes, cases, ref):
            x, y, z = (np.zeros(m.n_nodes) for _ in range(3))
            L.check("download", fe.lib.bf_state_download(fe.ctx, L.dp(x), L.dp(y), L.dp(z), None))
            n = m.n_v
            assert np.array_equal(x[:n], r[0][:n]) and np.array_equal(z[:n], r[2][:n])
        b.close()
    finally:
        for fe in fes:
            fe.close()
