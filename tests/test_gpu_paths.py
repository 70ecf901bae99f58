"""GPU coverage beyond the flat membrane: irregular Loop patches (ellipsoid dome), the
activation variants of ElemActDefGrad, per-element activation delays, the LV apex-cap CST
branch, UpdateElementThickness and the error codes of the C ABI."""
import copy
import ctypes as C

import numpy as np
import pytest
import torch

from biofem_amd import _lib as L
from biofem_amd import configs
from biofem_amd.fem import ActStrainFE
from biofem_amd.mesh import build_patch, flat_membrane
from oracle import bf_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def check_all(m, p, x, y, z, tol_exact=0.0, tol_fast=1e-12, tol_mixed=1e-5, omp=0):
    """Plain FP64 against the oracle bit for bit; the staged paths (with their hand-over of
    ill-conditioned elements to the plain kernel) within 1e-12 / 1e-5 of the oracle."""
    ref, st, c33 = O.finternal_act(m, p, x, y, z, want_c33=True, omp_threads=omp)
    fe = ActStrainFE(m, p)
    try:
        got = fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64)
        flags = fe.elem_flags()
        th = fe.thickness()
        e = rel(got, ref)
        print(f"  plain {e:.3e} bitwise {np.array_equal(got, ref)}; clamp-regime elements "
              f"{int((flags & L.FLAG_CLAMP).astype(bool).sum())} of {m.n_elmt}", flush=True)
        assert e <= tol_exact
        if tol_exact == 0.0:
            assert np.array_equal(fe.c33(), c33)
        # Each staged path twice: the first evaluation runs the narrow hand-over solve, the
        # second the solve the list length read back after the first picks (the wide kernel
        # for short lists); both add the CC terms in the reference's order, so the two
        # results are identical.
        fast = fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64_FAST)
        nf = fe.fallback_count()
        assert np.array_equal(fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64_FAST), fast)
        assert fe.fallback_count() == nf
        # the (element, qp) pairs handed over run in the reference order: bit for bit the
        # oracle's C33
        pf = check_handed_qps(fe, nf, c33)
        mixed = fe.finternal_act(x, y, z, precision=L.BF_PREC_MIXED)
        nm = fe.fallback_count()
        assert np.array_equal(fe.finternal_act(x, y, z, precision=L.BF_PREC_MIXED), mixed)
        assert fe.fallback_count() == nm
        pm = check_handed_qps(fe, nm, c33)
        fe.set_fallback(False)
        fast0 = fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64_FAST)
        mixed0 = fe.finternal_act(x, y, z, precision=L.BF_PREC_MIXED)
        waves = 4 * torch.cuda.get_device_properties(0).multi_processor_count
        print(f"  fast {rel(fast, ref):.3e} ({nf} elements / {pf} qps handed over, split "
              f"{split_lanes(pf, waves)}) mixed {rel(mixed, ref):.3e} ({nm} / {pm} handed "
              f"over, split {split_lanes(pm, waves)}); without hand-over: fast "
              f"{rel(fast0, ref):.3e} mixed {rel(mixed0, ref):.3e}", flush=True)
        assert rel(fast, ref) <= tol_fast
        assert rel(mixed, ref) <= tol_mixed
        return th, c33
    finally:
        fe.close()


def test_ellipsoid_small_irregular():
    m, p, xyz = configs.ellipsoid_shell(levels=3)
    assert (m.ire == 1).sum() > 0
    print("case ellipsoid levels=3", m.n_elmt, flush=True)
    check_all(m, p, *xyz)


def test_ellipsoid_full_bitwise():
    m, p, xyz = configs.ellipsoid_shell(levels=6)
    print("case ellipsoid levels=6", m.n_elmt, flush=True)
    check_all(m, p, *xyz, omp=16)


def test_membrane_large_staged_paths():
    m, p, xyz = configs.membrane_large()
    print("case membrane 256x192 staged paths", flush=True)
    check_all(m, p, *xyz, omp=16)


def split_lanes(n, n_waves):
    """Lanes per (element, qp) pair the hand-over solve picks for a list of n pairs: the
    largest of 60, 30, 15 (the wide kernel), 8, 6, 4, 2 whose pairs fit one round of the
    persistent grid (64 // k groups per wave); else 1."""
    for k in (60, 30, 15, 8, 6, 4, 2):
        if -(-n // (64 // k)) <= n_waves:
            return k
    return 1


def check_handed_qps(fe, n_elements, c33):
    """After a staged evaluation: the elements flagged as handed over are the counted ones,
    each has at least one (element, qp) pair solved in the reference order and no other
    element has one, and those pairs -- all points of an element whose CC3333 guard tripped,
    which is redone in that order -- hold the oracle's C33 bit for bit.  Returns the pair
    count."""
    flags = fe.elem_flags()
    handed = (flags & L.FLAG_FALLBACK) != 0
    assert int(handed.sum()) == n_elements
    hq = fe.handed_qps()
    assert not hq[~handed].any()
    assert hq[handed].any(axis=1).all()
    n_pairs = int(hq.sum())
    hq[handed & ((flags & L.FLAG_CC3333) != 0)] = True
    assert np.array_equal(fe.c33()[hq], c33[hq])
    return n_pairs


@pytest.mark.parametrize("sigma", [4e-4, 5.5e-4, 6.5e-4, 7e-4, 1e-3, 1.2e-3, 1.5e-3])
def test_membrane_handover_sizes(sigma):
    """Hand-over lists from a few dozen to 19k of the 98 304 elements (current z noise
    N(0, sigma), seed 11), for the FP64 and the mixed path: on a 256-CU device the solve
    splits each (element, qp) pair over 60, 30, 15 (the wide kernel), 8, 6, 4, 2 or 1 lanes by
    the list length, and every handed-over pair stays bit for bit the oracle's."""
    m, p, (x, y, z) = configs.membrane_large()
    rng = np.random.default_rng(11)
    z = m.z0.copy()
    z[: m.n_v] += rng.normal(0.0, sigma, m.n_v)
    m.apply_ghosts(x, y, z)
    print(f"case membrane 256x192 z noise {sigma}", flush=True)
    check_all(m, p, x, y, z, omp=16)


@pytest.mark.parametrize("mesh", ["membrane", "ellipsoid"])
@pytest.mark.parametrize("n_qp", [1, 2, 4, 5])
def test_quadrature_rules(n_qp, mesh):
    """SetGaussianQuadrature's rules other than the 3-point one (active_strain.c:184-222):
    the device's table and the lane groups of n_qp lanes (64 / n_qp elements per wave)
    against the oracle, on a displaced flat membrane and on the irregular patches of the
    small ellipsoid."""
    # On the shell the reference order's own rounding is ~1e-12 of the force (the staged
    # chain is within 1e-13 of an extended-precision evaluation, the oracle 1.4e-12 away:
    # test_staged_cpu.py::test_reference_rounding_on_ellipsoid), so the staged FP64 path's
    # stated tolerance there is 5e-12.
    if mesh == "membrane":
        m, p, _ = configs.membrane_small()
        x, y, z = configs.smooth_displacement(m, 0.01)
        tol_fast = 1e-12
    else:
        m, p, (x, y, z) = configs.ellipsoid_shell(levels=3)
        tol_fast = 5e-12
    p.n_qp = n_qp
    print(f"case {mesh} n_qp {n_qp}", flush=True)
    check_all(m, p, x, y, z, tol_fast=tol_fast)


@pytest.mark.parametrize("mode", ["isotropic", "manufactured", "ramp3_delay", "ramp1",
                                  "ramp2", "ramp2_delay"])
def test_activation_variants(mode):
    m, p, _ = configs.membrane_small()
    x, y, z = configs.smooth_displacement(m, 0.01)
    tol_exact = 0.0
    if mode == "isotropic":
        p.cart_fib_act = 0
    elif mode == "manufactured":
        p.cart_fib_act, p.manufactured = 0, 1
        p.manufactured_gamma0, p.manufactured_T, p.ti, p.dt = 0.1, 1.0, 30, 0.01
    elif mode == "ramp3_delay":
        p.ramp_type, p.ti, p.dt = 3, 20, 0.01
        m.activation_delay = np.linspace(0.0, 0.1, m.n_elmt)
        tol_exact = 1e-13  # device exp() against glibc exp()
    elif mode == "ramp1":
        p.ramp_type, p.ti, p.dt, p.tisteps = 1, 25, 0.01, 100
    elif mode == "ramp2":   # half-sine ramp, active_strain.c:1175-1178
        p.ramp_type, p.ti, p.dt, p.tisteps = 2, 40, 0.01, 100
    else:                   # the same with per-element delays, evaluated on the device
        p.ramp_type, p.ti, p.dt, p.tisteps = 2, 40, 0.01, 100
        m.activation_delay = np.linspace(0.0, 0.2, m.n_elmt)
        tol_exact = 1e-13  # device sin() against glibc sin()
    print("case", mode, flush=True)
    check_all(m, p, x, y, z, tol_exact=tol_exact, tol_fast=1e-12)


def test_thickness_matches_reference_formula():
    m, p, (x, y, z) = configs.membrane_small()
    th, c33 = check_all(m, p, x, y, z)
    mean = c33.mean(axis=1)
    np.testing.assert_allclose(th, p.h0 * np.sqrt(mean), rtol=1e-15)


def test_lv_cap_branch():
    """Elements >= n_elmt_base use the flat-triangle (CST) membrane branch
    (active_strain.c:448-472, :1807-1854)."""
    m = flat_membrane(12, 9, 1.0, 0.75, jitter=1e-3, seed=0)
    base = m.n_elmt - 20
    m.n_elmt_base = base
    m.patch = build_patch(m.n_v, m.n_nodes, m.n_elmt, m.tri, m.ire, m.irv, m.val, 16,
                          n_elmt_base=base)
    p = configs.membrane_small()[1]
    x, y, z = configs.smooth_displacement(m, 0.01)
    print("case lv cap", flush=True)
    check_all(m, p, x, y, z, tol_fast=1e-12)


def test_error_codes():
    lib = L.lib()
    m, p, (x, y, z) = configs.membrane_small()
    ctx = C.c_void_p()
    L.check("create", lib.bf_ctx_create(0, C.byref(ctx)))
    try:
        assert lib.bf_force_eval(ctx, L.dp(x), L.dp(y), L.dp(z), L.dp(np.zeros(3 * m.n_nodes)), 0, 0) == L.BF_ERR_STATE
        bad = copy.deepcopy(m)
        bad.patch[5, 4] = L.SENTINEL
        d = bad.desc()
        assert lib.bf_mesh_upload(ctx, C.byref(d)) == L.BF_ERR_ELEMENT
        d = m.desc()
        L.check("upload", lib.bf_mesh_upload(ctx, C.byref(d)))
        q = p.to_c()
        q.n_qp = 6
        assert lib.bf_params_set(ctx, C.byref(q)) == L.BF_ERR_ARG
        q = p.to_c()
        L.check("params", lib.bf_params_set(ctx, C.byref(q)))
        assert lib.bf_force_eval(ctx, L.dp(x), L.dp(y), L.dp(z), L.dp(np.zeros(3 * m.n_nodes)), 0, 9) == L.BF_ERR_ARG
        assert lib.bf_ctx_create(99, C.byref(C.c_void_p())) == L.BF_ERR_ARG
    finally:
        lib.bf_ctx_destroy(ctx)


def _noisy_reference(nx, amp, seed):
    """Membrane whose REFERENCE surface carries N(0, amp) z-noise, evaluated at x = X: the
    unnormalised a3,alpha of the noisy reference drives some elements through the
    CC3333 guard of ModElemC33 (active_strain.c:1674-1681) or to non-finite forces."""
    from biofem_amd import mesh as MM
    m, p, _ = configs.membrane_large(nx, nx * 3 // 4)
    rng = np.random.default_rng(seed)
    m.z0 = m.z0.copy()
    m.z0[: m.n_v] += rng.normal(0.0, amp, m.n_v)
    m.apply_ghosts(m.x0, m.y0, m.z0)
    m.dA0 = MM.area0(m.tri[: m.n_elmt], m.x0, m.y0, m.z0)
    return m, p, (m.x0.copy(), m.y0.copy(), m.z0.copy())


def test_cc3333_guard_case():
    """One element trips the guard: qps after the failing one keep their C33 in that
    iteration, exactly as the reference's early return does."""
    m, p, (x, y, z) = _noisy_reference(128, 3e-3, 2)
    print("case guard 128x96 amp 3e-3 seed 2", flush=True)
    ref, st, c33 = O.finternal_act(m, p, x, y, z, want_c33=True)
    assert st[2] == 1 and np.isfinite(ref).all()
    fe = ActStrainFE(m, p)
    try:
        got = fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64)
        assert np.array_equal(got, ref)
        assert np.array_equal(fe.c33(), c33)
        e, code, warn = fe.last_error()
        assert warn == st[2]
        assert fe.elem_flags()[st[1]] & 1
        # the staged paths hand the element over; the finish redoes it with coupled qps
        # (its own lanes for n_qp <= 4): its C33 is the reference's bit for bit
        g = st[1]
        fast = fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64_FAST)
        assert np.array_equal(fe.c33()[g], c33[g])
        assert fe.last_error()[2] >= 1 and fe.elem_flags()[g] & 1
        mixed = fe.finternal_act(x, y, z, precision=L.BF_PREC_MIXED)
        assert np.array_equal(fe.c33()[g], c33[g])
        assert rel(fast, ref) <= 1e-12 and rel(mixed, ref) <= 1e-5
    finally:
        fe.close()


def test_nonfinite_case_reported():
    """Elements whose force is not finite: the plain path reproduces the same values
    (NaN in the same places), returns BF_ERR_ELEMENT and reports BF_ELEM_NONFINITE with the
    first such element."""
    m, p, (x, y, z) = _noisy_reference(128, 3e-3, 0)
    print("case non-finite 128x96 amp 3e-3 seed 0", flush=True)
    ref, st, _ = O.finternal_act(m, p, x, y, z, want_c33=True)
    assert not np.isfinite(ref).all()
    fe = ActStrainFE(m, p)
    try:
        got = np.zeros(3 * m.n_nodes)
        rc = fe.lib.bf_force_eval(fe.ctx, L.dp(x), L.dp(y), L.dp(z), L.dp(got), 0, L.BF_PREC_FP64)
        assert rc == L.BF_ERR_ELEMENT  # forces are still written (PETSC_ERR_FP class)
        assert np.array_equal(got, ref, equal_nan=True)
        e, code, warn = fe.last_error()
        assert code == L.BF_ELEM_NONFINITE and e >= 0
        # the library counts elements, the oracle counts guard trips (element x iteration)
        assert 0 < warn <= st[2]
        assert int((fe.elem_flags() & L.FLAG_CC3333).astype(bool).sum()) == warn
    finally:
        fe.close()


def test_status_after_eval_on_caller_stream():
    """bf_last_error, bf_elem_flags_download and bf_c33_download read the status of an
    evaluation that bf_force_eval_device ran on a caller's stream, without the caller
    synchronising that stream first: the reads are ordered after it."""
    m, p, (x, y, z) = _noisy_reference(128, 3e-3, 0)
    print("case non-finite 128x96 on a caller stream", flush=True)
    ref, st, c33 = O.finternal_act(m, p, x, y, z, want_c33=True)
    fe = ActStrainFE(m, p)
    try:
        # a finite evaluation first, on the context's stream (the same mesh flattened:
        # finite forces, one CC3333 trip): its status has no non-finite element
        xs, ys, zs = x, y, np.zeros_like(z)
        fe.finternal_act(xs, ys, zs, precision=L.BF_PREC_FP64)
        assert fe.last_error()[1] != L.BF_ELEM_NONFINITE
        dev = torch.device("cuda:0")
        bx, by, bz = (torch.as_tensor(a, device=dev) for a in (x, y, z))
        f = torch.zeros(3 * m.n_nodes, dtype=torch.float64, device=dev)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        vp = C.c_void_p
        for _ in range(2):
            with torch.cuda.stream(side):
                rc = fe.lib.bf_force_eval_device(fe.ctx, vp(bx.data_ptr()), vp(by.data_ptr()),
                                                 vp(bz.data_ptr()), vp(f.data_ptr()), 0,
                                                 L.BF_PREC_FP64, vp(side.cuda_stream))
            assert rc == L.BF_OK
            e, code, warn = fe.last_error()      # no side.synchronize() before the read
            assert code == L.BF_ELEM_NONFINITE and e >= 0
            assert int((fe.elem_flags() & L.FLAG_CC3333).astype(bool).sum()) == warn
            assert np.array_equal(fe.c33(), c33)
            # and back to a finite state on the context's stream
            fe.finternal_act(xs, ys, zs, precision=L.BF_PREC_FP64)
            assert fe.last_error()[1] != L.BF_ELEM_NONFINITE
        side.synchronize()
        assert np.array_equal(f.cpu().numpy(), ref, equal_nan=True)
    finally:
        fe.close()


def test_timing_events_only_while_timing_is_on():
    """The timing events cost device time between the kernels, so an evaluation records them
    only while timing is on: bf_last_timing answers BF_ERR_STATE before, the split totals
    after bf_timing_enable(1); the same for a batch's bf_batch_last_ms."""
    from biofem_amd.batch import Batch, upload_coords
    m, p, (x, y, z) = configs.membrane_small()
    print("case timing events 32x24", flush=True)
    fe = ActStrainFE(m, p)
    bt = None
    try:
        fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64_FAST)
        a, b = C.c_float(), C.c_float()
        assert fe.lib.bf_last_timing(fe.ctx, C.byref(a), C.byref(b)) == L.BF_ERR_STATE
        L.check("t", fe.lib.bf_timing_enable(fe.ctx, 1))
        fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64_FAST)
        assert fe.lib.bf_last_timing(fe.ctx, C.byref(a), C.byref(b)) == L.BF_OK
        assert a.value > 0.0 and b.value > 0.0
        L.check("t", fe.lib.bf_timing_enable(fe.ctx, 0))
        upload_coords(fe, x, y, z)
        bt = Batch([fe])
        bt.eval(L.BF_PREC_FP64_FAST, graph=False)
        ms = C.c_float()
        assert fe.lib.bf_batch_last_ms(bt.h, C.byref(ms)) == L.BF_ERR_STATE
        bt.timing(True)
        bt.eval(L.BF_PREC_FP64_FAST, graph=False)
        assert bt.last_ms() > 0.0
    finally:
        if bt is not None:
            bt.close()
        fe.close()


def test_status_words_start_reset_every_step():
    """The staged paths' node pass hands the status words over and resets them for the next
    step (no reset launch then): repeated evaluations report the same counts, not sums, also
    after a batch step of the same context wrote its own."""
    from biofem_amd.batch import Batch, upload_coords
    m, p, (x, y, z) = _noisy_reference(128, 3e-3, 2)
    print("case status reset guard 128x96 amp 3e-3 seed 2", flush=True)
    fe = ActStrainFE(m, p)
    bt = None
    try:
        seen = []
        for prec in (L.BF_PREC_FP64_FAST, L.BF_PREC_FP64_FAST, L.BF_PREC_MIXED,
                     L.BF_PREC_FP64_FAST):
            fe.finternal_act(x, y, z, precision=prec)
            seen.append((prec, fe.last_error(), fe.fallback_count()))
        assert seen[0] == seen[1] == seen[3]
        assert seen[0][1][2] >= 1  # the guard trip is reported once per step
        upload_coords(fe, x, y, z)
        bt = Batch([fe])
        bt.eval(L.BF_PREC_FP64_FAST, graph=True)
        bt.eval(L.BF_PREC_FP64_FAST, graph=True)
        bt.sync()  # the batch's stream is not ordered with the context's
        fe.finternal_act(x, y, z, precision=L.BF_PREC_FP64_FAST)
        assert (L.BF_PREC_FP64_FAST, fe.last_error(), fe.fallback_count()) == seen[0]
    finally:
        if bt is not None:
            bt.close()
        fe.close()
