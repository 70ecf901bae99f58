"""Served sweep: requests admitted by HBM footprint into continuous batches; a case's
result is the same alone and in the mix (bit for bit: its arithmetic does not depend on
which cases share its launches)."""
import numpy as np
import pytest

from biofem_amd import _lib as L
from biofem_amd.serve import MeshCache, Server, draw_sweep, run_alone

pytestmark = pytest.mark.gpu


def test_sweep_overlap_admission_and_identity():
    specs = draw_sweep(n_cases=12, rate=400.0, seed=11, elmt_lo=2000, elmt_hi=12000,
                       steps_lo=20, steps_hi=60)
    cache = MeshCache()
    foot = sorted(cache.get(s).footprint_bytes(state=True) for s in specs)
    budget = sum(foot[-4:])          # room for about four cases: later arrivals must queue
    print("case sweep of 12, budget", budget, flush=True)
    srv = Server(hbm_budget=budget, steps_per_round=8, keep_results=True)
    rep = srv.run(specs)
    s = rep.summary()
    print(s, flush=True)
    assert s["cases"] == len(specs)
    assert rep.max_active >= 2                   # work of several cases shared launches
    assert rep.peak_footprint <= budget
    assert s["queue_wait_p99_s"] > 0.0           # admission by memory made someone wait
    by_id = {c.spec.cid: c for c in rep.cases}
    for spec in specs[:4]:
        alone = run_alone(spec, steps_per_round=5)
        mixed = by_id[spec.cid].result
        n = by_id[spec.cid].mesh.n_v
        assert np.isfinite(mixed).all()
        assert np.array_equal(alone[:n], mixed[:n]), spec.cid


def test_async_rounds_contract_and_identity():
    """bf_batch_step_async / bf_batch_step_wait: two rounds in flight at most, the case set
    frozen while one is, and rounds enqueued ahead give the synchronous steps' state bit for
    bit."""
    from biofem_amd.batch import Batch
    from biofem_amd.serve import build_case, case_result
    specs = draw_sweep(n_cases=2, rate=100.0, seed=3, elmt_lo=2000, elmt_hi=6000,
                       steps_lo=20, steps_hi=20)
    cache = MeshCache()
    print("case async rounds of", [s.cid for s in specs], flush=True)
    cases = [build_case(s, cache) for s in specs]
    b = Batch([c.fe for c in cases])
    try:
        b.step_async(L.BF_PREC_MIXED, 4, 0)
        b.step_async(L.BF_PREC_MIXED, 4, 1)
        with pytest.raises(L.BfError):        # round 0 not waited for
            b.step_async(L.BF_PREC_MIXED, 4, 2)
        with pytest.raises(L.BfError):        # rounds in flight
            b.set_cases([c.fe for c in cases])
        for r in (0, 1):
            ms, nonfinite, _ = b.step_wait(r, len(cases))
            assert ms > 0.0 and (nonfinite == -1).all()
        with pytest.raises(L.BfError):        # already waited for
            b.step_wait(1, len(cases))
        for r in range(2, 5):                 # 20 steps in all, one round kept ahead
            b.step_async(L.BF_PREC_MIXED, 4, r)
            b.step_wait(r - 1, len(cases)) if r > 2 else None
        b.step_wait(4, len(cases))
        got = [case_result(c) for c in cases]
    finally:
        b.close()
        for c in cases:
            c.fe.close()
    for spec, g in zip(specs, got):
        alone = run_alone(spec, steps_per_round=4)
        assert np.array_equal(alone, g), spec.cid


def test_sweep_full_size_range():
    """Requests at both ends of the sweep's range: a 2 k and a 300 k-triangle flat shell and
    a 3 k and a 184 k-triangle (level 6) ellipsoid, thin shells, served together; every case
    completes, the percentiles cover all of them, and the largest cases give the same state
    alone as in the mix."""
    from biofem_amd.serve import CaseSpec
    specs = [CaseSpec(0, "flat", 2100, 1.2, 0.006, 0.12, 40, 0.00, seed=100),
             CaseSpec(1, "ellipsoid", 184320, 0.9, 0.005, 0.1, 40, 0.01, seed=101),
             CaseSpec(2, "flat", 300000, 1.5, 0.006, 0.15, 30, 0.02, seed=102),
             CaseSpec(3, "ellipsoid", 3000, 1.8, 0.012, 0.18, 40, 0.03, seed=103)]
    print("case sweep full size range", flush=True)
    srv = Server(hbm_budget=16 << 30, steps_per_round=10, keep_results=True)
    rep = srv.run(specs)
    s = rep.summary()
    print({k: v for k, v in s.items() if k != "failed_cases"}, flush=True)
    assert s["cases"] == 4 and s["failed"] == 0 and s["completed"] == 4
    assert s["largest_ellipsoid"] == 184320 and s["largest_flat"] >= 270000
    assert s["latency_p99_s"] is not None and s["latency_p99_s"] == s["latency_completed_p99_s"]
    by_id = {c.spec.cid: c for c in rep.cases}
    for cid in (1, 2):
        spec = specs[cid]
        alone = run_alone(spec, steps_per_round=10)
        mixed = by_id[cid].result
        n = by_id[cid].mesh.n_v
        assert np.isfinite(mixed).all()
        assert np.array_equal(alone[:n], mixed[:n]), cid


@pytest.mark.parametrize("prec", [L.BF_PREC_FP64, L.BF_PREC_MIXED])
def test_thick_fine_shell_diverges_in_the_reference_arithmetic(prec):
    """A level-6 ellipsoid (184 k triangles) with h0 = 0.02 is outside the reference
    formulation's range: the unnormalised a3,alpha of ElemUpdateG (active_strain.c:745-762,
    :940-957) gives off-mid-surface points a normal metric term of order h0 / h at the
    extraordinary-vertex patches, whose C33 collapses already at rest (the oracle's own
    trajectory goes non-finite: tests/test_model_limits.py).  Stepped in the reference's
    operation order (BF_PREC_FP64) as in FP32-mixed, the case is retired with its reason
    and counted in the report, its latency inside the percentiles."""
    from biofem_amd.serve import CaseSpec
    specs = [CaseSpec(0, "ellipsoid", 184320, 1.0, 0.02, 0.15, 60, 0.0, seed=7),
             CaseSpec(1, "flat", 2100, 1.0, 0.01, 0.1, 60, 0.0, seed=8)]
    print(f"case thick level-6 ellipsoid, precision {prec}", flush=True)
    srv = Server(hbm_budget=16 << 30, steps_per_round=10, precision=prec)
    rep = srv.run(specs)
    s = rep.summary()
    print(s, flush=True)
    assert s["cases"] == 2 and s["failed"] == 1 and s["completed"] == 1
    assert s["failed_cases"][0]["cid"] == 0 and "non-finite" in s["failed_cases"][0]["reason"]
    assert s["latency_p50_s"] is not None and s["latency_max_s"] is not None
