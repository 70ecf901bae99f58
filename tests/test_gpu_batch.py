"""Batched evaluation (one persistent element launch over the work list of many meshes,
hipGraph-captured steps, launch-shape autotuner): every case's forces equal its own
single-mesh evaluation bit for bit, in every precision and launch shape."""
import numpy as np
import pytest

from biofem_amd import _lib as L
from biofem_amd import configs
from biofem_amd.batch import Batch, upload_coords
from biofem_amd.fem import ActStrainFE
from oracle import bf_oracle as O

pytestmark = pytest.mark.gpu

PRECS = [L.BF_PREC_FP64, L.BF_PREC_FP64_FAST, L.BF_PREC_MIXED]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _cases(n, seed=2, nx_hi=40, mixed_qp=False):
    cases = configs.batch_membranes(n=n, seed=seed, nx_lo=8, nx_hi=nx_hi)
    if mixed_qp:  # lane groups sized by the largest n_qp; smaller cases leave lanes idle
        cases[0][1].n_qp, cases[1][1].n_qp, cases[2][1].n_qp = 1, 5, 2
        cases[3][1].c33_iters = 5
    return cases


def _single(cases, prec):
    out = []
    for m, p, (x, y, z) in cases:
        fe = ActStrainFE(m, p)
        out.append(fe.finternal_act(x, y, z, precision=prec))
        fe.close()
    return out


def _resident(cases):
    fes = []
    for m, p, (x, y, z) in cases:
        fe = ActStrainFE(m, p)
        upload_coords(fe, x, y, z)
        fes.append(fe)
    return fes


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("mixed_qp", [False, True])
def test_batch_equals_single(prec, mixed_qp):
    cases = _cases(7, mixed_qp=mixed_qp)
    print(f"case batch of {len(cases)} prec {prec} mixed_qp {mixed_qp}", flush=True)
    ref = _single(cases, prec)
    fes = _resident(cases)
    b = Batch(fes)
    try:
        for graph in (False, True, True):
            b.eval(prec, ghosts=False, graph=graph)
            b.sync()
            for i in range(len(fes)):
                assert np.array_equal(b.forces(i), ref[i]), (graph, i)
    finally:
        b.close()
        for fe in fes:
            fe.close()


@pytest.mark.parametrize("prec", [L.BF_PREC_FP64_FAST, L.BF_PREC_MIXED])
def test_batch_handover_lists_equal_single(prec):
    """Rough current states (z noise of one mesh spacing h) make every case hand about a
    tenth of its elements to the reference order: the batch's concatenated hand-over list
    gives every case's forces bit for bit as its own evaluation."""
    rng = np.random.default_rng(9)
    cases = []
    for m, p, _ in _cases(7):
        h = 1.0 / np.sqrt(m.n_elmt / 1.5)
        x, y, z = m.x0.copy(), m.y0.copy(), m.z0.copy()
        z[: m.n_v] += rng.normal(0.0, h, m.n_v)
        m.apply_ghosts(x, y, z)
        cases.append((m, p, (x, y, z)))
    print(f"case batch of {len(cases)} rough states prec {prec}", flush=True)
    ref = _single(cases, prec)
    fes = _resident(cases)
    b = Batch(fes)
    try:
        for graph in (False, True):
            b.eval(prec, ghosts=False, graph=graph)
            b.sync()
            assert b.fallback_count() > 0
            for i in range(len(fes)):
                assert np.array_equal(b.forces(i), ref[i]), (graph, i)
    finally:
        b.close()
        for fe in fes:
            fe.close()


def test_batch_plain_matches_oracle():
    cases = _cases(3)
    print("case batch plain vs oracle", flush=True)
    fes = _resident(cases)
    b = Batch(fes)
    try:
        b.eval(L.BF_PREC_FP64, graph=True)
        b.sync()
        for i, (m, p, (x, y, z)) in enumerate(cases):
            ref, st, _ = O.finternal_act(m, p, x, y, z)
            assert np.array_equal(b.forces(i), ref)
    finally:
        b.close()
        for fe in fes:
            fe.close()


def test_batch_ghosts_and_graph_replays():
    """Ghost rows are recomputed by the captured step; replays interleaved with eager
    single-mesh work give the same forces every time."""
    cases = _cases(5)
    print("case batch ghosts + replays", flush=True)
    ref = _single(cases, L.BF_PREC_MIXED)
    fes = _resident(cases)
    for fe, (m, p, (x, y, z)) in zip(fes, cases):
        xg, yg, zg = x.copy(), y.copy(), z.copy()
        xg[m.n_v:] = yg[m.n_v:] = zg[m.n_v:] = 123.0   # stale ghosts
        upload_coords(fe, xg, yg, zg)
    b = Batch(fes)
    other = ActStrainFE(*cases[0][:2])
    try:
        for rep in range(4):
            b.eval(L.BF_PREC_MIXED, ghosts=True, graph=True)
            b.sync()
            for i in range(len(fes)):
                assert np.array_equal(b.forces(i), ref[i]), (rep, i)
            other.finternal_act(*cases[0][2], precision=L.BF_PREC_FP64)
    finally:
        other.close()
        b.close()
        for fe in fes:
            fe.close()


def test_batch_tuner_and_launch_shapes():
    """Launch shapes of the single-pass staged kernel: the tile size (threads per block) sets
    the node sums' order, so each shape is compared bit for bit with the single-mesh
    evaluation under the same tiling, and every shape within 1e-13 of the default one; the
    grid (persistent or one tile per workgroup) changes nothing."""
    cases = _cases(6, nx_hi=64)
    print("case batch tuner", flush=True)
    ref = _single(cases, L.BF_PREC_FP64_FAST)
    fes = _resident(cases)
    b = Batch(fes)

    def own_single():
        return [fe.finternal_act(*c[2], precision=L.BF_PREC_FP64_FAST) for fe, c in zip(fes, cases)]

    try:
        for passes, block, persistent in ((1, 64, False), (1, 64, True), (8, 128, True),
                                          (3, 256, True), (1, 256, False)):
            b.set_launch(passes, block, 0, persistent)
            b.eval(L.BF_PREC_FP64_FAST, graph=False)
            b.sync()
            got = [b.forces(i) for i in range(len(fes))]
            same = own_single()
            for i in range(len(fes)):
                assert np.array_equal(got[i], same[i]), (block, persistent, i)
                assert rel(got[i], ref[i]) <= 1e-13, (block, i)
        best = b.tune(L.BF_PREC_FP64_FAST, reps=2)
        print("tuned", best, b.info(), flush=True)
        assert best["ms"] > 0
        b.eval(L.BF_PREC_FP64_FAST, graph=True)
        b.sync()
        got = [b.forces(i) for i in range(len(fes))]
        same = own_single()
        for i in range(len(fes)):
            assert np.array_equal(got[i], same[i])
        # a smaller set of cases on the same batch
        b.set_cases(fes[:2])
        b.eval(L.BF_PREC_FP64_FAST, graph=True)
        b.sync()
        for i in range(2):
            assert np.array_equal(b.forces(i), got[i])
    finally:
        b.close()
        for fe in fes:
            fe.close()


@pytest.mark.parametrize("prec", [L.BF_PREC_FP64, L.BF_PREC_MIXED])
def test_batch_case_status_flags_diverged_case(prec):
    """A case whose coordinates hold a NaN reports its first non-finite element through
    bf_batch_case_status; the other cases report none, and the flags clear once the case's
    coordinates are finite again."""
    cases = _cases(4)
    print(f"case batch status, prec {prec}", flush=True)
    fes = _resident(cases)
    m, p, (x, y, z) = cases[2]
    bad = x.copy()
    bad[5] = np.nan
    b = Batch(fes)
    try:
        upload_coords(fes[2], bad, y, z)
        b.eval(prec, ghosts=False, graph=False)
        b.sync()
        nf, cc = b.case_status()
        assert nf[2] >= 0 and (nf[[0, 1, 3]] == -1).all(), nf
        flags = fes[2].elem_flags()
        assert flags[nf[2]] & L.FLAG_NONFINITE
        upload_coords(fes[2], x, y, z)
        b.eval(prec, ghosts=False, graph=False)
        b.sync()
        nf, cc = b.case_status()
        assert (nf == -1).all() and (cc == -1).all(), (nf, cc)
    finally:
        b.close()
        for fe in fes:
            fe.close()


def test_batch_config_full_size_vs_oracle():
    """The 64-membrane batch at its stated sizes (nx log-uniform in [16, 256], seed 2;
    1.02 M triangles), autotuned launch shape, hipGraph replays: every case's forces
    against the OpenMP oracle -- plain FP64 bit for bit, staged FP64 within 1e-12 and
    FP32-mixed within 1e-5 (2-norm relative, per case)."""
    cases = configs.batch_membranes()
    n_el = sum(m.n_elmt for m, _, _ in cases)
    print(f"case batch of {len(cases)} membranes, {n_el} triangles: oracle", flush=True)
    refs = [O.finternal_act(m, p, x, y, z, omp_threads=16)[0] for m, p, (x, y, z) in cases]
    fes = _resident(cases)
    b = Batch(fes)
    try:
        best = b.tune(L.BF_PREC_MIXED, reps=2)
        print("  tuned", best, flush=True)
        worst = {}
        for prec in PRECS:
            for rep in range(2):           # capture, then a replay
                b.eval(prec, ghosts=True, graph=True)
                b.sync()
                errs = [rel(b.forces(i), refs[i]) for i in range(len(cases))]
                if prec == L.BF_PREC_FP64:
                    assert all(np.array_equal(b.forces(i), refs[i]) for i in range(len(cases)))
                worst[prec] = max(errs)
            print(f"  precision {prec}: worst case rel err {worst[prec]:.3e}, handed over "
                  f"{b.fallback_count()}", flush=True)
        assert worst[L.BF_PREC_FP64_FAST] <= 1e-12
        assert worst[L.BF_PREC_MIXED] <= 1e-5
    finally:
        b.close()
        for fe in fes:
            fe.close()
