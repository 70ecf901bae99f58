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
