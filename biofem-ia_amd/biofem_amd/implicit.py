"""The reference's implicit time step, on device: Newmark inertia, matrix-free Newton-Krylov.

Each step of the reference (src/main.c:421-575) solves R(x) = 0 with PETSc SNES over a
matrix-free (finite-difference) Jacobian, R = Fint(x) - Fext + Fdyn(x) from FormFunctionFEM
(src/main.c:757-1032; FDynamic's Newmark inertia src/main.c:1157-1300, lumped element
masses rho h0 A0 / 3 and Rayleigh mass damping, src/material.c:100-140; a fixed degree of
freedom's residual replaced by x - x0, src/external.c:1706-1790), then updates velocity
and acceleration (xAccVel, src/main.c:1367-1420) and shifts x -> xn -> xnm1.

PETSc is not available, so the solver is this module's own: Newton with backtracking on
||R||, restarted GMRES with classical Gram-Schmidt and no preconditioner, and the Walker-
Pernice differencing parameter, with PETSc's default tolerances.  The force is the
library's (bf_ghost_update_device + bf_force_eval_device on torch-owned device buffers,
on a stream ordered against torch's current one); every other operation is FP64 torch.
The solver functions take any residual callable, so they run on the CPU as well.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L

SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)   # PETSC_SQRT_MACHINE_EPSILON


def gmres(matvec, b: torch.Tensor, rtol: float = 1e-5, atol: float = 1e-50,
          restart: int = 30, max_it: int = 10000):
    """x with ||b - A x|| <= max(rtol ||b||, atol) (KSPGMRES defaults, x0 = 0, no
    preconditioner).  Returns (x, iterations, converged, residual norm)."""
    n = b.numel()
    x = torch.zeros_like(b)
    bnorm = float(torch.linalg.vector_norm(b))
    tol = max(rtol * bnorm, atol)
    if bnorm <= tol:
        return x, 0, True, bnorm
    its = 0
    r = b.clone()
    beta = bnorm
    while its < max_it:
        m = min(restart, max_it - its)
        V = torch.zeros((m + 1, n), dtype=b.dtype, device=b.device)
        H = np.zeros((m + 1, m))
        V[0] = r / beta
        cs, sn = np.zeros(m), np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        k_done = 0
        for k in range(m):
            w = matvec(V[k])
            # classical Gram-Schmidt in two matrix products; one transfer to the host
            h = V[: k + 1] @ w
            w = w - h @ V[: k + 1]
            hv = torch.cat([h, torch.linalg.vector_norm(w).view(1)]).double().cpu().numpy()
            hn = float(hv[-1])
            H[: k + 2, k] = hv
            for i in range(k):  # earlier Givens rotations
                t = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
                H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
                H[i, k] = t
            den = math.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = (H[k, k] / den, H[k + 1, k] / den) if den > 0 else (1.0, 0.0)
            H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            its += 1
            k_done = k + 1
            if abs(g[k + 1]) <= tol or hn == 0.0:
                break
            V[k + 1] = w / hn
        y = np.linalg.solve(np.triu(H[:k_done, :k_done]), g[:k_done]) if k_done else np.zeros(0)
        x = x + torch.as_tensor(y, dtype=b.dtype, device=b.device) @ V[:k_done]
        r = b - matvec(x)
        beta = float(torch.linalg.vector_norm(r))
        if beta <= tol:
            return x, its, True, beta
    return x, its, False, beta


@dataclass
class SolveInfo:
    converged: bool
    reason: str
    iterations: int
    linear_iterations: int
    fnorms: list = field(default_factory=list)


def jfnk(residual, U0: torch.Tensor, rtol: float = 1e-8, atol: float = 1e-50,
         stol: float = 1e-8, max_it: int = 50, ksp_rtol: float = 1e-5, ksp_restart: int = 30,
         ksp_max_it: int = 10000, mffd_err: float = SQRT_EPS, ls_alpha: float = 1e-4,
         ls_min_lambda: float = 1e-12):
    """Newton's method on residual(U) = 0 (SNESNEWTONLS with backtracking, the default
    convergence tests) with Jacobian-vector products (R(U + h a) - R(U)) / h,
    h = mffd_err sqrt(1 + ||U||) / ||a|| (MatMFFD 'wp').  Returns (U, SolveInfo)."""
    U = U0.clone()
    F = residual(U)
    fnorm = float(torch.linalg.vector_norm(F))
    info = SolveInfo(False, "", 0, 0, [fnorm])
    if not math.isfinite(fnorm):
        info.reason = "diverged: residual not finite"
        return U, info
    if fnorm < atol:
        info.converged, info.reason = True, "converged: ||F|| < atol"
        return U, info
    f0 = fnorm
    for it in range(1, max_it + 1):
        unorm = float(torch.linalg.vector_norm(U))

        c = mffd_err * math.sqrt(1.0 + unorm)

        def jv(a, U=U, F=F, c=c):
            # h stays on the device (no host round trip per product); a = 0 gives
            # R(U) - F = 0 and a zero product, as a skipped one would
            h = c / torch.linalg.vector_norm(a).clamp_min(1e-300)
            return (residual(U + h * a) - F) / h

        d, kits, _, _ = gmres(jv, -F, rtol=ksp_rtol, restart=ksp_restart, max_it=ksp_max_it)
        info.linear_iterations += kits
        lam = 1.0
        while True:
            Ut = U + lam * d
            Ft = residual(Ut)
            ft = float(torch.linalg.vector_norm(Ft))
            if math.isfinite(ft) and ft * ft <= (1.0 - 2.0 * ls_alpha * lam) * fnorm * fnorm:
                break
            lam *= 0.5
            if lam < ls_min_lambda:
                info.iterations = it
                info.reason = "diverged: line search failed"
                return U, info
        snorm = lam * float(torch.linalg.vector_norm(d))
        U, F, fnorm = Ut, Ft, ft
        info.iterations = it
        info.fnorms.append(fnorm)
        if fnorm < atol:
            info.converged, info.reason = True, "converged: ||F|| < atol"
            return U, info
        if fnorm <= rtol * f0:
            info.converged, info.reason = True, "converged: ||F|| <= rtol ||F0||"
            return U, info
        if snorm < stol * float(torch.linalg.vector_norm(U)):
            info.converged, info.reason = True, "converged: ||dx|| < stol ||x||"
            return U, info
    info.reason = "diverged: max_it"
    return U, info


class NewmarkJFNK:
    """Implicit Newmark steps (timeinteg 0, constant average acceleration for
    newmark_gamma = 0.5; beta = (gamma + 1/2)^2 / 4) of the state x, xn, xd, xdd, each a
    torch FP64 vector [3 n_nodes] interleaved as the reference's Vecs.

    internal_force(U) -> Fint [3 n_nodes] with ghost rows zero; `for_fe` builds it from an
    ActStrainFE (GlobalGhost + the library's force on device).  mass [n_nodes]: assembled
    lumped masses (lumped_masses); fixed [3 n_nodes] bool: degrees of freedom held at x0.
    damping / dampfactor: -damping and -dampfactor; the mass-proportional damping matrix
    C = dampfactor M enters the residual only when damping is set, as FDynamic applies Damp
    only under -damping (src/main.c:1257, :1326), otherwise C = 0.
    element_thickness() -> [n_elmt]: UpdateElementThickness of the last force evaluation,
    called once per converged step as the reference does (src/main.c:497-503) and kept in
    `thickness` (default: internal_force.thickness when the callable has one)."""

    def __init__(self, internal_force, x0: torch.Tensor, mass: torch.Tensor, dt: float,
                 n_v: int | None = None, fixed: torch.Tensor | None = None,
                 fext: torch.Tensor | None = None, newmark_gamma: float = 0.5,
                 damping: bool = False, dampfactor: float = 0.0, element_thickness=None,
                 **solver):
        self.fint = internal_force
        self.element_thickness = (element_thickness if element_thickness is not None
                                  else getattr(internal_force, "thickness", None))
        self.thickness = None
        self.x0 = x0.clone()
        n3 = x0.numel()
        self.n_v = n3 // 3 if n_v is None else n_v
        self.dt = dt
        g = newmark_gamma
        b = 0.25 * (g + 0.5) * (g + 0.5)
        self.M1, self.M2, self.M3 = 1.0 / (b * dt * dt), 1.0 / (b * dt), 1.0 / (2 * b) - 1.0
        self.C1, self.C2, self.C3 = g / (b * dt), g / b - 1.0, dt * (g / (2.0 * b) - 1.0)
        self.V2, self.V3 = (1.0 - g) * dt, g * dt          # xAccVel's C2, C3
        self.M = mass.repeat_interleave(3).to(x0)
        self.C = (dampfactor if damping else 0.0) * self.M
        self.fixed = (torch.zeros(n3, dtype=torch.bool, device=x0.device) if fixed is None
                      else fixed.to(x0.device))
        self.ghost = torch.zeros(n3, dtype=torch.bool, device=x0.device)
        self.ghost[3 * self.n_v:] = True
        self.fext = torch.zeros_like(x0) if fext is None else fext.to(x0)
        # free rows: R = Fint + a U - b with a = M M1 + C C1 and, per step,
        # b = Fext + M (M1 xn + M2 xd + M3 xdd) + C (C1 xn + C2 xd + C3 xdd); ghost rows of a
        # and b are zero, so R's ghost rows are Fint's (zero); fixed rows: U - x0
        self.a = torch.where(self.ghost, 0.0, self.M * self.M1 + self.C * self.C1)
        self.fixed_rows = self.fixed & ~self.ghost
        self.b = torch.zeros_like(x0)
        self.x = x0.clone()
        self.xn = x0.clone()
        self.xnm1 = x0.clone()
        self.xd = torch.zeros_like(x0)
        self.xdd = torch.zeros_like(x0)
        self.solver = solver
        self.history: list = []

    def fdyn(self, U: torch.Tensor) -> torch.Tensor:
        """FDynamic, timeinteg 0 (src/main.c:1268-1272) with assembled lumped masses."""
        facc = self.M * (self.M1 * self.xn + self.M2 * self.xd + self.M3 * self.xdd)
        fvel = self.C * (self.C1 * self.xn + self.C2 * self.xd + self.C3 * self.xdd)
        return self.M * self.M1 * U + self.C * self.C1 * U - facc - fvel

    def residual(self, U: torch.Tensor) -> torch.Tensor:
        """FormFunctionFEM's residual at U (the b of the current step)."""
        R = torch.addcmul(self.fint(U), self.a, U).sub_(self.b)
        return torch.where(self.fixed_rows, U - self.x0, R)

    def step(self) -> SolveInfo:
        facc = self.M * (self.M1 * self.xn + self.M2 * self.xd + self.M3 * self.xdd)
        fvel = self.C * (self.C1 * self.xn + self.C2 * self.xd + self.C3 * self.xdd)
        self.b = torch.where(self.ghost, 0.0, self.fext + facc + fvel)
        U, info = jfnk(self.residual, self.x, **self.solver)
        if info.converged:
            self.x = U
            # the last residual was evaluated at U (the accepted line-search point)
            if self.element_thickness is not None:
                self.thickness = self.element_thickness()
        # xAccVel (src/main.c:1367-1420), then the shift of src/main.c:563-564
        xdd_old = self.xdd
        self.xdd = self.M1 * (self.x - self.xn) - self.M2 * self.xd - self.M3 * self.xdd
        self.xd = self.xd + self.V2 * xdd_old + self.V3 * self.xdd
        self.xnm1, self.xn = self.xn, self.x.clone()
        self.history.append(info)
        return info


def lumped_masses(m, rho: float, h0: float) -> np.ndarray:
    """Node masses [n_nodes]: each real triangle gives rho h0 dA0 / 3 to each corner
    (Mass, src/material.c:100-120, assembled); ghost nodes 0."""
    out = np.zeros(m.n_nodes)
    w = rho * h0 * np.asarray(m.dA0) / 3.0
    T = np.asarray(m.tri[: m.n_elmt])
    for k in range(3):
        np.add.at(out, T[:, k], w)
    return out


def for_fe(fe, precision: int = L.BF_PREC_FP64_FAST, device: str = "cuda"):
    """internal_force(U) over an ActStrainFE: U [3 n_nodes] -> coordinates (SoA device
    buffers), GlobalGhost, the library's force on device, ghost rows zeroed.  Launch
    errors raise; a force that is not finite shows in the residual norm, which jfnk
    reports as a diverged step."""
    m = fe.mesh
    n = m.n_nodes
    bxyz = torch.zeros((3, n), dtype=torch.float64, device=device)  # SoA rows x, y, z
    bx, by, bz = bxyz[0], bxyz[1], bxyz[2]
    f = torch.zeros(3 * n, dtype=torch.float64, device=device)
    lib = fe.lib
    vp = C.c_void_p

    # the library's launches go to a stream of their own, ordered against the caller's
    # current stream both ways (the default stream has no handle to pass: a NULL stream
    # would send the launches to the context's stream, unordered with torch's work)
    side = torch.cuda.Stream(device=device)

    def internal_force(U: torch.Tensor) -> torch.Tensor:
        cur = torch.cuda.current_stream(device)
        bxyz.copy_(U.view(-1, 3).t())
        side.wait_stream(cur)
        s = vp(side.cuda_stream)
        L.check("bf_ghost_update_device", lib.bf_ghost_update_device(
            fe.ctx, vp(bx.data_ptr()), vp(by.data_ptr()), vp(bz.data_ptr()), s))
        rc = lib.bf_force_eval_device(fe.ctx, vp(bx.data_ptr()), vp(by.data_ptr()),
                                      vp(bz.data_ptr()), vp(f.data_ptr()), 0, precision, s)
        L.check("bf_force_eval_device", rc)
        cur.wait_stream(side)
        return f.clone()      # accumulate = 0: ghost rows are written as zeros

    def thickness() -> np.ndarray:
        side.synchronize()    # the library reads its C33 state on its own stream
        return fe.thickness()

    internal_force.thickness = thickness
    return internal_force
