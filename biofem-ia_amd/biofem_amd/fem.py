"""Python mirror of the reference's FE object around the force call.

``ActStrainFE`` holds what FInternalAct (src/active_strain.c:2226-2267) reads -- mesh,
options, globals -- and evaluates it on the GPU through the C ABI.  Options use the
reference's names (-num_gaussian_quad_points, -C33_subitr_nums, -cart_fib_act, -mu,
-bulk_modulus, -muscle_act_gamma, -gamma_ramp_type, ...; src/active_strain.c:94-108).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib as L
from .mesh import Mesh


@dataclass
class ActParams:
    n_qp: int = 3                 # -num_gaussian_quad_points
    c33_iters: int = 3            # -C33_subitr_nums
    cart_fib_act: int = 1         # -cart_fib_act
    manufactured: int = 0
    h0: float = 0.01
    mu: float = 1.0 / (2.0 * 1.3)        # E / (2 (1 + nu)), E = 1, nu = 0.3
    K: float = 1.0 / (3.0 * (1 - 0.6))   # E / (3 (1 - 2 nu))
    gamma: float = 0.1            # -muscle_act_gamma
    ramp_type: int = 0            # -gamma_ramp_type
    gamma_T: float = 0.0
    tau_up: float = 0.0
    tau_down: float = 0.0
    apd: float = 0.0
    pulse_t0: float = 0.0
    pulse_sigma: float = 1.0
    manufactured_gamma0: float = 0.0
    manufactured_T: float = 1.0
    ti: int = 0
    tisteps: int = 1
    dt: float = 1.0

    @staticmethod
    def from_young(E: float, nu: float, **kw) -> "ActParams":
        return ActParams(mu=E / (2.0 * (1.0 + nu)), K=E / (3.0 * (1.0 - 2.0 * nu)), **kw)

    def to_c(self) -> L.BfParams:
        p = L.BfParams()
        for f in fields(self):
            setattr(p, f.name, getattr(self, f.name))
        return p


class ActStrainFE:
    """One mesh resident on one GPU; ``finternal_act`` is the FInternalAct drop-in."""

    def __init__(self, mesh: Mesh, params: ActParams, device: int = 0):
        self.lib = L.lib()
        self.mesh = mesh
        self.params = params
        ctx = C.c_void_p()
        L.check("bf_ctx_create", self.lib.bf_ctx_create(device, C.byref(ctx)))
        self.ctx = ctx
        d = mesh.desc()
        L.check("bf_mesh_upload", self.lib.bf_mesh_upload(self.ctx, C.byref(d)))
        self.set_params(params)

    def set_params(self, params: ActParams) -> None:
        self.params = params
        p = params.to_c()
        L.check("bf_params_set", self.lib.bf_params_set(self.ctx, C.byref(p)))

    def finternal_act(self, x, y, z, fint=None, accumulate=False, precision=L.BF_PREC_FP64):
        n = self.mesh.n_nodes
        x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
        assert len(x) == len(y) == len(z) == n
        if fint is None:
            fint = np.zeros(3 * n)
            accumulate = False
        rc = self.lib.bf_force_eval(self.ctx, L.dp(x), L.dp(y), L.dp(z), L.dp(fint),
                                    int(accumulate), int(precision))
        L.check("bf_force_eval", rc)
        return fint

    def ghost_update(self, x, y, z) -> None:
        L.check("bf_ghost_update", self.lib.bf_ghost_update(self.ctx, L.dp(x), L.dp(y), L.dp(z)))

    def last_error(self):
        e, c_, w = C.c_int(), C.c_int(), C.c_int()
        L.check("bf_last_error", self.lib.bf_last_error(self.ctx, C.byref(e), C.byref(c_), C.byref(w)))
        return e.value, c_.value, w.value

    def last_timing(self):
        a, b = C.c_float(), C.c_float()
        L.check("bf_last_timing", self.lib.bf_last_timing(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def c33(self) -> np.ndarray:
        out = np.zeros(self.mesh.n_elmt * self.params.n_qp)
        L.check("bf_c33_download", self.lib.bf_c33_download(self.ctx, L.dp(out)))
        return out.reshape(self.mesh.n_elmt, self.params.n_qp)

    def fallback_count(self) -> int:
        n = C.c_int()
        L.check("bf_last_fallback_count", self.lib.bf_last_fallback_count(self.ctx, C.byref(n)))
        return n.value

    def handed_qps(self) -> np.ndarray:
        """[n_elmt, n_qp] bool: the (element, qp) pairs the last staged evaluation solved in
        the reference's operation order (bf_last_handed_qps); elements whose CC3333 guard
        tripped (FLAG_CC3333) were redone in that order with all their points."""
        mask = np.zeros(self.mesh.n_elmt, np.uint8)
        n = C.c_int()
        L.check("bf_last_handed_qps", self.lib.bf_last_handed_qps(
            self.ctx, mask.ctypes.data_as(C.POINTER(C.c_ubyte)), C.byref(n)))
        bits = (mask[:, None] >> np.arange(self.params.n_qp)[None, :]) & 1
        out = bits.astype(bool)
        assert int(out.sum()) == n.value
        return out

    def set_fallback(self, on: bool) -> None:
        L.check("bf_set_fallback", self.lib.bf_set_fallback(self.ctx, int(on)))

    def elem_flags(self) -> np.ndarray:
        out = np.zeros(self.mesh.n_elmt, np.uint8)
        L.check("bf_elem_flags_download",
                self.lib.bf_elem_flags_download(self.ctx, out.ctypes.data_as(C.POINTER(C.c_ubyte))))
        return out

    def thickness(self) -> np.ndarray:
        out = np.zeros(self.mesh.n_elmt)
        L.check("bf_update_thickness", self.lib.bf_update_thickness(self.ctx, L.dp(out)))
        return out

    def device_buffers(self):
        ptrs = [C.c_void_p() for _ in range(4)]
        L.check("bf_device_buffers", self.lib.bf_device_buffers(self.ctx, *[C.byref(p) for p in ptrs]))
        return [p.value for p in ptrs]

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.bf_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
