"""Batched force evaluation of many meshes (cases) in the same launches.

Each case is an ActStrainFE (one mesh + options with resident device state); a Batch
evaluates all of them per step with one persistent element kernel driven by a device work
list of element tiles, optionally replayed from a captured hipGraph, with a launch-shape
autotuner (bf_batch_* in include/biofem_hip.h).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib as L
from .fem import ActStrainFE


class Batch:
    def __init__(self, cases: Sequence[ActStrainFE] = (), device: int = 0):
        self.lib = L.lib()
        h = C.c_void_p()
        L.check("bf_batch_create", self.lib.bf_batch_create(device, C.byref(h)))
        self.h = h
        self.cases: list = []
        if cases:
            self.set_cases(cases)

    def set_cases(self, cases: Sequence[ActStrainFE]) -> None:
        self.cases = list(cases)
        arr = (C.c_void_p * max(len(self.cases), 1))(*[c.ctx.value for c in self.cases])
        L.check("bf_batch_set_cases", self.lib.bf_batch_set_cases(self.h, arr, len(self.cases)))

    def refresh(self) -> None:
        L.check("bf_batch_refresh", self.lib.bf_batch_refresh(self.h))

    def eval(self, precision=L.BF_PREC_MIXED, ghosts=False, graph=True) -> None:
        L.check("bf_batch_eval", self.lib.bf_batch_eval(self.h, int(precision), int(ghosts), int(graph)))

    def sync(self) -> None:
        L.check("bf_batch_sync", self.lib.bf_batch_sync(self.h))

    def timing(self, on: bool) -> None:
        """HIP events around every later eval/step call (bf_batch_timing_enable), for last_ms;
        each record costs a few microseconds of device time."""
        L.check("bf_batch_timing_enable", self.lib.bf_batch_timing_enable(self.h, int(bool(on))))

    def last_ms(self) -> float:
        ms = C.c_float()
        L.check("bf_batch_last_ms", self.lib.bf_batch_last_ms(self.h, C.byref(ms)))
        return ms.value

    def step_async(self, precision, steps: int, round_no: int, graph: bool = True) -> None:
        """Enqueue round `round_no` (bf_batch_step + its per-case status) and return."""
        L.check("bf_batch_step_async", self.lib.bf_batch_step_async(
            self.h, int(precision), int(steps), int(graph), int(round_no)))

    def step_wait(self, round_no: int, n_cases: int):
        """Wait for round `round_no`: (GPU ms, first non-finite element per case, first
        CC3333 trip per case), -1 for none, in the case order the round ran with."""
        ms = C.c_float()
        nf = np.full(max(n_cases, 1), -1, np.int32)
        cc = np.full(max(n_cases, 1), -1, np.int32)
        ip = C.POINTER(C.c_int)
        L.check("bf_batch_step_wait", self.lib.bf_batch_step_wait(
            self.h, int(round_no), C.byref(ms), nf.ctypes.data_as(ip), cc.ctypes.data_as(ip)))
        return ms.value, nf[:n_cases], cc[:n_cases]

    def case_status(self):
        """Per case of the batch: (first non-finite element, first CC3333 trip) of the last
        step, -1 for none."""
        n = len(self.cases)
        nf = np.full(n, -1, np.int32)
        cc = np.full(n, -1, np.int32)
        if n:
            ip = C.POINTER(C.c_int)
            L.check("bf_batch_case_status",
                    self.lib.bf_batch_case_status(self.h, nf.ctypes.data_as(ip), cc.ctypes.data_as(ip)))
        return nf, cc

    def fallback_count(self) -> int:
        n = C.c_int()
        L.check("bf_batch_last_fallback_count", self.lib.bf_batch_last_fallback_count(self.h, C.byref(n)))
        return n.value

    def info(self):
        nc, ne, nt, cap = C.c_int(), C.c_longlong(), C.c_int(), C.c_int()
        L.check("bf_batch_info", self.lib.bf_batch_info(self.h, C.byref(nc), C.byref(ne), C.byref(nt), C.byref(cap)))
        return dict(cases=nc.value, elements=ne.value, tiles=nt.value, tile_cap=cap.value)

    def set_launch(self, passes: int, block: int, blocks: int = 0, persistent: bool = True) -> None:
        L.check("bf_batch_set_launch", self.lib.bf_batch_set_launch(self.h, passes, block, blocks, int(persistent)))

    def launch(self):
        v = [C.c_int() for _ in range(4)]
        L.check("bf_batch_get_launch", self.lib.bf_batch_get_launch(self.h, *[C.byref(x) for x in v]))
        return dict(passes=v[0].value, block=v[1].value, blocks=v[2].value, persistent=v[3].value)

    def tune(self, precision=L.BF_PREC_MIXED, reps: int = 3):
        v = [C.c_int() for _ in range(4)]
        ms = C.c_float()
        L.check("bf_batch_tune", self.lib.bf_batch_tune(self.h, int(precision), reps,
                                                        *[C.byref(x) for x in v], C.byref(ms)))
        return dict(passes=v[0].value, block=v[1].value, blocks=v[2].value,
                    persistent=v[3].value, ms=ms.value)

    def forces(self, i: int) -> np.ndarray:
        fe = self.cases[i]
        out = np.zeros(3 * fe.mesh.n_nodes)
        L.check("bf_fint_download", self.lib.bf_fint_download(fe.ctx, L.dp(out)))
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.bf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def upload_coords(fe: ActStrainFE, x, y, z) -> None:
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    L.check("bf_coords_upload", fe.lib.bf_coords_upload(fe.ctx, L.dp(x), L.dp(y), L.dp(z)))
