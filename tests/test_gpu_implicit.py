"""The implicit Newmark / matrix-free Newton-Krylov step on the device (biofem_amd.implicit)
against a CPU restatement of the reference's residual on the oracle's force: after every
step, FormFunctionFEM's residual -- Fint (oracle) + FDynamic's per-element Newmark inertia,
fixed rows x - x0, ghost rows 0 -- is at the Newton tolerance, and xAccVel's update of
velocity and acceleration holds.  Parity with PETSc's SNES itself is unpinned (PETSc is not
available); the test pins the equations the solver solves."""
import numpy as np
import pytest
import torch

from biofem_amd import _lib as L
from biofem_amd import configs
from biofem_amd.fem import ActStrainFE
from biofem_amd.implicit import NewmarkJFNK, for_fe, lumped_masses
from oracle import bf_oracle as O

pytestmark = pytest.mark.gpu


def _oracle_residual(m, p, U, xn, xd, xdd, x0, fixed, rho, dt, gam=0.5, dampfactor=0.0):
    x, y, z = U[0::3].copy(), U[1::3].copy(), U[2::3].copy()
    m.apply_ghosts(x, y, z)
    fint, _, _ = O.finternal_act(m, p, x, y, z, want_c33=True)
    fint = np.asarray(fint, dtype=np.float64).copy()
    fint[3 * m.n_v:] = 0.0
    beta = 0.25 * (gam + 0.5) ** 2
    M1, M2, M3 = 1 / (beta * dt * dt), 1 / (beta * dt), 1 / (2 * beta) - 1
    C1, C2, C3 = gam / (beta * dt), gam / beta - 1, dt * (gam / (2 * beta) - 1)
    fdyn = np.zeros_like(U)
    T = np.asarray(m.tri[: m.n_elmt])
    Me = rho * p.h0 * np.asarray(m.dA0) / 3.0        # Mass (src/material.c:100-120)
    Ce = dampfactor * Me
    for k in range(3):
        for d in range(3):
            i = 3 * T[:, k] + d
            fd = (Me * M1 * U[i] + Ce * C1 * U[i] - Me * (M1 * xn[i] + M2 * xd[i] + M3 * xdd[i])
                  - Ce * (C1 * xn[i] + C2 * xd[i] + C3 * xdd[i]))
            np.add.at(fdyn, i, fd)
    R = fint + fdyn
    R[fixed] = U[fixed] - x0[fixed]
    R[3 * m.n_v:] = 0.0
    return R


@pytest.mark.parametrize("dampfactor", [0.0, 5.0])
def test_implicit_steps_solve_the_reference_residual(dampfactor):
    m, p, (x, y, z) = configs.membrane_small()
    print(f"case membrane 32x24 implicit Newmark, dampfactor {dampfactor}", flush=True)
    rho, dt = 1.0, 1e-2
    fe = ActStrainFE(m, p)
    try:
        x0 = np.stack([m.x0, m.y0, m.z0], 1).reshape(-1)
        fixed = np.zeros(3 * m.n_nodes, bool)
        clamp = np.where(np.isclose(m.x0[: m.n_v], 0.0))[0]
        for d in range(3):
            fixed[3 * clamp + d] = True
        mass = lumped_masses(m, rho, p.h0)
        dev = torch.device("cuda")
        st = NewmarkJFNK(for_fe(fe, L.BF_PREC_FP64_FAST), torch.as_tensor(x0, device=dev),
                         torch.as_tensor(mass, device=dev), dt, n_v=m.n_v,
                         fixed=torch.as_tensor(fixed), damping=dampfactor != 0.0,
                         dampfactor=dampfactor)
        # the library's force at the reference state against the oracle's
        f_dev = st.fint(st.x).cpu().numpy()
        f_ref, _, _ = O.finternal_act(m, p, *(a.copy() for a in (m.x0, m.y0, m.z0)), want_c33=True)
        f_ref = np.asarray(f_ref).copy()
        f_ref[3 * m.n_v:] = 0.0
        assert np.linalg.norm(f_dev - f_ref) <= 1e-12 * np.linalg.norm(f_ref)
        for s in range(3):
            xn, xd, xdd = (a.cpu().numpy().copy() for a in (st.xn, st.xd, st.xdd))
            r0 = _oracle_residual(m, p, st.x.cpu().numpy(), xn, xd, xdd, x0, fixed, rho, dt,
                                  dampfactor=dampfactor)
            info = st.step()
            print(f"  step {s}: {info.reason}, {info.iterations} Newton / "
                  f"{info.linear_iterations} GMRES iterations, |R| {info.fnorms[0]:.3e} -> "
                  f"{info.fnorms[-1]:.3e}", flush=True)
            assert info.converged, info
            U = st.x.cpu().numpy()
            r = _oracle_residual(m, p, U, xn, xd, xdd, x0, fixed, rho, dt, dampfactor=dampfactor)
            assert np.linalg.norm(r) <= 1e-6 * np.linalg.norm(r0), (np.linalg.norm(r), np.linalg.norm(r0))
            # xAccVel (src/main.c:1367-1420), gamma 1/2, beta 1/4
            beta, gam = 0.25, 0.5
            a_new = (U - xn) / (beta * dt * dt) - xd / (beta * dt) - (1 / (2 * beta) - 1) * xdd
            v_new = xd + (1 - gam) * dt * xdd + gam * dt * a_new
            assert np.allclose(st.xdd.cpu().numpy(), a_new, rtol=1e-12, atol=1e-9)
            assert np.allclose(st.xd.cpu().numpy(), v_new, rtol=1e-12, atol=1e-11)
            # UpdateElementThickness after the converged step (src/active_strain.c:2198-2224)
            x, y, z = U[0::3].copy(), U[1::3].copy(), U[2::3].copy()
            m.apply_ghosts(x, y, z)
            _, _, c33 = O.finternal_act(m, p, x, y, z, want_c33=True)
            th_ref = p.h0 * np.sqrt(c33.mean(axis=1))
            assert np.abs(st.thickness - th_ref).max() <= 1e-11 * p.h0
        assert np.abs(U[fixed] - x0[fixed]).max() <= 1e-10
        assert np.abs(U - x0).max() > 1e-6      # the active membrane moved
    finally:
        fe.close()


def test_implicit_100_steps_config1():
    """The 32 x 24 active membrane (x = 0 edge clamped, rho 1, dt 1e-2, no damping) for
    100 implicit Newmark steps, as the reference's main loop runs them (src/main.c:420-575):
    every step converges, and after every step FormFunctionFEM's residual restated on the
    oracle's force is at the Newton tolerance relative to the step's starting residual."""
    m, p, _ = configs.membrane_small()
    print("case membrane 32x24, 100 implicit Newmark steps", flush=True)
    rho, dt = 1.0, 1e-2
    fe = ActStrainFE(m, p)
    try:
        x0 = np.stack([m.x0, m.y0, m.z0], 1).reshape(-1)
        fixed = np.zeros(3 * m.n_nodes, bool)
        clamp = np.where(np.isclose(m.x0[: m.n_v], 0.0))[0]
        for d in range(3):
            fixed[3 * clamp + d] = True
        mass = lumped_masses(m, rho, p.h0)
        dev = torch.device("cuda")
        st = NewmarkJFNK(for_fe(fe, L.BF_PREC_FP64_FAST), torch.as_tensor(x0, device=dev),
                         torch.as_tensor(mass, device=dev), dt, n_v=m.n_v,
                         fixed=torch.as_tensor(fixed))
        worst = 0.0
        for s in range(100):
            xn, xd, xdd = (a.cpu().numpy().copy() for a in (st.xn, st.xd, st.xdd))
            r0 = _oracle_residual(m, p, st.x.cpu().numpy(), xn, xd, xdd, x0, fixed, rho, dt)
            info = st.step()
            assert info.converged, (s, info)
            r = _oracle_residual(m, p, st.x.cpu().numpy(), xn, xd, xdd, x0, fixed, rho, dt)
            ratio = np.linalg.norm(r) / np.linalg.norm(r0)
            worst = max(worst, ratio)
            assert ratio <= 1e-6, (s, ratio)
            if s % 20 == 19:
                print(f"  step {s + 1}: {info.iterations} Newton / {info.linear_iterations} "
                      f"GMRES, oracle residual ratio {ratio:.2e} (worst {worst:.2e})",
                      flush=True)
        U = st.x.cpu().numpy()
        assert np.abs(U[fixed] - x0[fixed]).max() <= 1e-10
        assert np.isfinite(U).all() and np.abs(U - x0).max() > 1e-6
    finally:
        fe.close()
