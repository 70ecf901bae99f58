"""Served parameter sweep: the `biofem sweep` workflow as a request loop.

The reference's `biofem sweep` (bin/biofem:352-456) writes one case directory per parameter
combination and runs each case as its own MPI job.  Here a case is a request: it arrives
(Poisson process), is admitted when its device footprint fits the HBM budget and a slot is
free, keeps its state resident on the GPU (bf_ctx + bf_state_init), and advances in
continuous batches: every round runs the same hipGraph-captured step of ALL admitted cases
(one persistent element launch over the union of their element tiles), so requests that
arrive while others run are merged into the next round and their work overlaps on the
device.  A case leaves when its step count is reached; its latency is completion minus
arrival.  Each case's arithmetic does not depend on which others share its launches, so
its result is the same alone and in the mix (bit for bit).
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .batch import Batch, upload_coords
from .fem import ActParams, ActStrainFE
from .mesh import Mesh, ellipsoid_dome, flat_membrane, lumped_mass

# Ellipsoid subdivision levels a case may use: a request is served at the level nearest
# its triangle count (log scale).  Level 7 (737 280 triangles) lies above the sweep's
# largest request (400 000), so the largest served ellipsoid is level 6.
ELLIPSOID_SIZES = {3: 2880, 4: 11520, 5: 46080, 6: 184320}


@dataclass
class CaseSpec:
    cid: int
    kind: str            # "flat" or "ellipsoid"
    target_elmt: int     # requested triangle count (snapped to the generator's sizes)
    E_mod: float
    h0: float
    gamma: float
    steps: int
    arrival: float       # seconds after the start of the sweep
    seed: int = 0


def draw_sweep(n_cases: int = 2000, rate: float = 20.0, seed: int = 5, elmt_lo: int = 2000,
               elmt_hi: int = 400000, steps_lo: int = 50, steps_hi: int = 500) -> list:
    """Poisson arrivals (exponential gaps of mean 1/rate), kind flat/ellipsoid with equal
    probability, triangle count log-uniform in [elmt_lo, elmt_hi], materials and activation
    as in the batched workload (E U[0.5, 2], nu 0.3, h0 U[0.005, 0.02], gamma U[0.05, 0.2]),
    steps uniform in [steps_lo, steps_hi]; numpy default_rng(seed)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_cases))
    kind = rng.integers(0, 2, n_cases)
    ne = np.exp(rng.uniform(np.log(elmt_lo), np.log(elmt_hi), n_cases))
    E = rng.uniform(0.5, 2.0, n_cases)
    h0 = rng.uniform(0.005, 0.02, n_cases)
    gam = rng.uniform(0.05, 0.2, n_cases)
    steps = rng.integers(steps_lo, steps_hi + 1, n_cases)
    return [CaseSpec(i, "flat" if kind[i] == 0 else "ellipsoid", int(ne[i]), float(E[i]),
                     float(h0[i]), float(gam[i]), int(steps[i]), float(t[i]), seed=1000 + i)
            for i in range(n_cases)]


class MeshCache:
    """Reference meshes by resolution: flat nx snapped to a geometric grid of ratio 2^(1/4),
    ellipsoid subdivision level by nearest size (levels 3..5)."""

    def __init__(self):
        self.cache = {}

    def key(self, spec: CaseSpec):
        if spec.kind == "flat":
            nx = np.sqrt(spec.target_elmt / 1.5)
            k = int(round(4 * np.log2(max(nx, 8.0))))
            return ("flat", max(8, int(round(2 ** (k / 4.0)))))
        lv = min(ELLIPSOID_SIZES, key=lambda L_: abs(np.log(ELLIPSOID_SIZES[L_] / spec.target_elmt)))
        return ("ellipsoid", lv)

    def get(self, spec: CaseSpec) -> Mesh:
        k = self.key(spec)
        if k not in self.cache:
            if k[0] == "flat":
                nx = k[1]
                self.cache[k] = flat_membrane(nx, max(2, int(round(0.75 * nx))), 1.0, 0.75,
                                              jitter=1e-3, seed=0)
            else:
                self.cache[k] = ellipsoid_dome(levels=k[1])
        return self.cache[k]


@dataclass
class Case:
    spec: CaseSpec
    fe: ActStrainFE
    mesh: Mesh
    footprint: int
    remaining: int
    admitted: float = 0.0
    finished: float = 0.0
    result: np.ndarray | None = None
    failed: str | None = None   # why the case was retired before its last step


def stable_dt(m: Mesh, p: ActParams, rho: float = 1.0, safety: float = 0.1) -> float:
    """Explicit step below both the membrane limit h / c and the plate-bending limit
    ~0.7 h^2 / (h0 c) (bending waves: omega_max = (pi/h)^2 h0 c / sqrt(12)), with
    c = sqrt((K + 4/3 mu) / rho) and h the shortest edge-length scale sqrt(2 min dA0)."""
    h = float(np.sqrt(2.0 * m.dA0.min()))
    c = float(np.sqrt((p.K + 4.0 / 3.0 * p.mu) / rho))
    return safety * min(h / c, 0.7 * h * h / (p.h0 * c))


def build_case(spec: CaseSpec, cache: MeshCache, device: int = 0) -> Case:
    """Mesh (cached topology, per-case fibres), options, resident state at rest in the
    reference configuration; the clamped edge is x = 0 for flat membranes and the rim for
    ellipsoids.  Time step: stable_dt; damping 0.995 per step."""
    base = cache.get(spec)
    rng = np.random.default_rng(spec.seed)
    if spec.kind == "flat":
        ang = np.deg2rad(rng.uniform(-30.0, 30.0, base.n_elmt))
        fib = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], 1)
        m = dataclasses.replace(base, n_fib=fib)
        fixed = np.zeros(m.n_nodes, np.uint8)
        fixed[: m.n_v][np.isclose(m.x0[: m.n_v], 0.0)] = 1
    else:
        m = dataclasses.replace(base)
        fixed = np.zeros(m.n_nodes, np.uint8)
        fixed[np.asarray(m.extra["rim"])] = 1
    p = ActParams.from_young(spec.E_mod, 0.3, n_qp=3, c33_iters=3, cart_fib_act=1, h0=spec.h0,
                             gamma=spec.gamma, ramp_type=0)
    fe = ActStrainFE(m, p, device=device)
    upload_coords(fe, m.x0, m.y0, m.z0)
    mass = lumped_mass(m, 1.0, spec.h0)
    dt = stable_dt(m, p)
    L.check("bf_state_init", fe.lib.bf_state_init(fe.ctx, L.dp(mass), fixed.ctypes.data_as(
        L.C.POINTER(L.C.c_ubyte)), None, dt, 0.995))
    return Case(spec=spec, fe=fe, mesh=m, footprint=m.footprint_bytes(state=True),
                remaining=spec.steps)


def case_result(case: Case) -> np.ndarray:
    n = case.mesh.n_nodes
    x, y, z = np.zeros(n), np.zeros(n), np.zeros(n)
    L.check("bf_state_download", case.fe.lib.bf_state_download(case.fe.ctx, L.dp(x), L.dp(y),
                                                               L.dp(z), None))
    return np.stack([x, y, z], 1)


@dataclass
class SweepReport:
    cases: list = field(default_factory=list)
    wall: float = 0.0
    rounds: int = 0
    gpu_ms: float = 0.0
    max_active: int = 0
    peak_footprint: int = 0
    warmup: float = 0.0    # seconds of Server.warm before the first arrival

    def summary(self) -> dict:
        """Latency percentiles over every case that arrived: a case retired before its last
        step (its forces went non-finite) counts with the time from its arrival to its
        retirement, so failures stay inside p50/p99; the completed-only percentiles and
        the retired cases with their reasons are reported beside them."""
        def pct(a, q):
            return float(np.percentile(a, q)) if len(a) else None
        ok = [c for c in self.cases if c.failed is None]
        lat_all = np.array([c.finished - c.spec.arrival for c in self.cases])
        lat_ok = np.array([c.finished - c.spec.arrival for c in ok])
        wait = np.array([c.admitted - c.spec.arrival for c in self.cases])
        work = sum(c.mesh.n_elmt * (c.spec.steps - max(c.remaining, 0)) for c in self.cases)
        failed = [dict(cid=c.spec.cid, kind=c.spec.kind, triangles=c.mesh.n_elmt,
                       h0=round(c.spec.h0, 5), gamma=round(c.spec.gamma, 4),
                       steps=c.spec.steps, reason=c.failed)
                  for c in self.cases if c.failed is not None]
        big_ell = [c.mesh.n_elmt for c in self.cases if c.spec.kind == "ellipsoid"]
        big_flat = [c.mesh.n_elmt for c in self.cases if c.spec.kind == "flat"]
        return dict(
            cases=len(self.cases), completed=len(ok), failed=len(failed), wall_s=self.wall,
            throughput_cases_per_s=len(self.cases) / self.wall if self.wall else 0.0,
            elem_steps_per_s=work / self.wall if self.wall else 0.0,
            elem_qp_per_s_gpu=3 * work / (self.gpu_ms * 1e-3) if self.gpu_ms else 0.0,
            latency_p50_s=pct(lat_all, 50), latency_p99_s=pct(lat_all, 99),
            latency_max_s=float(lat_all.max()) if len(lat_all) else None,
            latency_completed_p50_s=pct(lat_ok, 50), latency_completed_p99_s=pct(lat_ok, 99),
            queue_wait_p50_s=pct(wait, 50), queue_wait_p99_s=pct(wait, 99),
            rounds=self.rounds, gpu_busy_s=self.gpu_ms * 1e-3, max_active=self.max_active,
            peak_footprint_gib=self.peak_footprint / 2**30, warmup_s=self.warmup,
            largest_ellipsoid=max(big_ell) if big_ell else 0,
            largest_flat=max(big_flat) if big_flat else 0,
            failed_cases=failed)


class Server:
    """Continuous-batching request loop over one GPU."""

    def __init__(self, hbm_budget: int, max_active: int = 256, steps_per_round: int = 10,
                 precision: int = L.BF_PREC_MIXED, graph: bool = True, device: int = 0,
                 keep_results: bool = False, log=None, warm_up: bool = True):
        self.budget = hbm_budget
        self.max_active = max_active
        self.k = steps_per_round
        self.precision = precision
        self.graph = graph
        self.device = device
        self.keep = keep_results
        self.cache = MeshCache()
        self.log = log or (lambda s: None)
        self.warm_up = warm_up

    def warm(self, specs) -> float:
        """Before the first arrival, as a server does before it takes traffic: build the
        reference mesh of every resolution the specs ask for, and step a short copy of the
        smallest case through the batch path (kernel code loaded, graph captured).  Returns
        the seconds spent.  Per-case work (fibres, context, upload) stays in each request's
        latency."""
        t = time.perf_counter()
        for s in specs:
            self.cache.get(s)
        if specs:
            small = min(specs, key=lambda s: self.cache.get(s).n_elmt)
            case = build_case(dataclasses.replace(small, steps=2), self.cache, self.device)
            batch = Batch(device=self.device)
            try:
                batch.set_cases([case.fe])
                for _ in range(2):
                    L.check("bf_batch_step", batch.lib.bf_batch_step(batch.h, self.precision, 1,
                                                                     int(self.graph)))
                batch.sync()
                batch.case_status()
            finally:
                batch.close()
                case.fe.close()
        return time.perf_counter() - t

    def run(self, specs, time_scale: float = 1.0) -> SweepReport:
        """Serve `specs` in arrival order; arrivals are replayed in wall-clock time divided
        by time_scale.  Rounds are enqueued one ahead of the one the host is waiting for
        (bf_batch_step_async), so the GPU does not idle while the host reads a round's
        status; the case set changes (an arrival admitted, a case finished or retired) only
        after the rounds in flight have been drained."""
        pending = sorted(specs, key=lambda s: s.arrival)
        rep = SweepReport()
        if self.warm_up:
            rep.warmup = self.warm(pending)
            self.log(f"warm-up {rep.warmup:.2f} s: {len(self.cache.cache)} mesh resolutions")
        batch = Batch(device=self.device)
        active: list = []
        inflight: deque = deque()   # (round, steps, [(case, steps done before the round)])
        used = 0
        rnd = 0
        t0 = time.perf_counter()
        t_log = 0.0
        i = 0
        changed = False

        def settle(entry):
            """Wait for a round; a case whose forces went non-finite in it is marked for
            retirement with the round's step window."""
            r, k, cases = entry
            ms, nonfinite, _ = batch.step_wait(r, len(cases))
            rep.gpu_ms += ms
            rep.rounds += 1
            for (c, start), e in zip(cases, nonfinite):
                if e >= 0 and c.failed is None:
                    c.failed = (f"non-finite force at element {int(e)} within steps "
                                f"{start + 1}..{start + k}")

        def fits(spec) -> bool:
            need = self.cache.get(spec).footprint_bytes(state=True)
            if need > self.budget:
                raise RuntimeError(f"case {spec.cid} needs {need} B, budget {self.budget} B")
            return used + need <= self.budget and len(active) < self.max_active

        try:
            while i < len(pending) or active or inflight:
                now = time.perf_counter() - t0
                arrived = i < len(pending) and pending[i].arrival / time_scale <= now
                if (not active or arrived and fits(pending[i])
                        or any(c.remaining <= 0 or c.failed for c in active)):
                    while inflight:
                        settle(inflight.popleft())
                    done = [c for c in active if c.remaining <= 0 or c.failed]
                    if done:
                        tnow = time.perf_counter() - t0
                        for c in done:
                            if self.keep and c.failed is None:
                                c.result = case_result(c)
                            if c.failed:
                                self.log(f"case {c.spec.cid} retired: {c.failed}")
                            c.finished = tnow
                            used -= c.footprint
                            c.fe.close()
                            rep.cases.append(c)
                        active = [c for c in active if c.remaining > 0 and not c.failed]
                        changed = True
                    # admission: arrived requests in order, while memory and slots allow
                    now = time.perf_counter() - t0
                    while (i < len(pending) and pending[i].arrival / time_scale <= now
                           and fits(pending[i])):
                        case = build_case(pending[i], self.cache, self.device)
                        case.admitted = time.perf_counter() - t0
                        active.append(case)
                        used += case.footprint
                        changed = True
                        i += 1
                    rep.max_active = max(rep.max_active, len(active))
                    rep.peak_footprint = max(rep.peak_footprint, used)
                    if not active:
                        if i >= len(pending):
                            break
                        nxt = pending[i].arrival / time_scale
                        time.sleep(max(0.0, nxt - (time.perf_counter() - t0)))
                        continue
                    if changed:
                        batch.set_cases([c.fe for c in active])
                        changed = False
                if now - t_log >= 5.0:
                    t_log = now
                    waiting = sum(1 for sp in pending[i:] if sp.arrival / time_scale <= now)
                    self.log(f"t={now:6.1f} s: {len(rep.cases)} done, {len(active)} active, "
                             f"{waiting} waiting, {len(pending) - i - waiting} yet to arrive, "
                             f"{used / 2**30:.2f} GiB resident")
                k = min(self.k, min(c.remaining for c in active))
                batch.step_async(self.precision, k, rnd, self.graph)
                inflight.append((rnd, k, [(c, c.spec.steps - c.remaining) for c in active]))
                for c in active:
                    c.remaining -= k
                rnd += 1
                if len(inflight) >= 2:   # one round queued behind the one that runs
                    settle(inflight.popleft())
            rep.wall = time.perf_counter() - t0
        finally:
            try:
                while inflight:
                    settle(inflight.popleft())
            finally:
                batch.close()
                for c in active:
                    c.fe.close()
        return rep


def run_alone(spec: CaseSpec, precision: int = L.BF_PREC_MIXED, steps_per_round: int = 10,
              graph: bool = True) -> np.ndarray:
    """The same case served by itself: its final coordinates."""
    cache = MeshCache()
    case = build_case(spec, cache)
    b = Batch([case.fe])
    try:
        left = spec.steps
        while left > 0:
            k = min(steps_per_round, left)
            L.check("bf_batch_step", b.lib.bf_batch_step(b.h, precision, k, int(graph)))
            left -= k
        b.sync()
        return case_result(case)
    finally:
        b.close()
        case.fe.close()
