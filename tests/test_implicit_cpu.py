"""The implicit step's solver pieces on the CPU (biofem_amd.implicit): restarted GMRES,
matrix-free Newton with backtracking, and the Newmark step with xAccVel's update, pinned
against direct linear algebra, scipy's root finder and a direct Newmark restatement on a
linear spring chain (where the constant-average-acceleration rule conserves energy)."""
import numpy as np
import torch
from scipy.optimize import fsolve

from biofem_amd.implicit import NewmarkJFNK, gmres, jfnk

T = lambda a: torch.as_tensor(a, dtype=torch.float64)


def test_gmres_solves_a_nonsymmetric_system():
    rng = np.random.default_rng(0)
    n = 60
    A = rng.standard_normal((n, n)) / np.sqrt(n) + 3.0 * np.eye(n)
    b = rng.standard_normal(n)
    x, its, ok, rn = gmres(lambda v: T(A) @ v, T(b), rtol=1e-12, restart=10)
    assert ok and its > 10                      # needed a restart
    assert np.linalg.norm(A @ x.numpy() - b) <= 1e-11 * np.linalg.norm(b)
    assert np.allclose(x.numpy(), np.linalg.solve(A, b), rtol=0, atol=1e-10)


def test_newton_krylov_on_a_nonlinear_system():
    rng = np.random.default_rng(1)
    n = 40
    A = np.diag(np.full(n, 4.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.5), -1)
    b = rng.standard_normal(n)
    F = lambda u: A @ u + 0.3 * u ** 3 - b
    U, info = jfnk(lambda u: T(F(u.numpy())), T(np.zeros(n)))
    assert info.converged and info.iterations < 15, info
    assert np.linalg.norm(F(U.numpy())) <= 1e-8 * np.linalg.norm(b)
    assert np.allclose(U.numpy(), fsolve(F, np.zeros(n), xtol=1e-13), atol=1e-9)


def _chain(n=12, k=50.0, m=0.7):
    K = np.diag(np.full(n, 2 * k)) - np.diag(np.full(n - 1, k), 1) - np.diag(np.full(n - 1, k), -1)
    K[-1, -1] = k
    return K, np.full(n, m)


def test_newmark_step_matches_direct_newmark_and_conserves_energy():
    """Springs along x only (3 dofs per node, y and z free and unloaded): the step's
    trajectory equals a direct solve of the same Newmark equations, and the energy
    1/2 v'Mv + 1/2 u'Ku is conserved by the average-acceleration rule."""
    n, dt, steps = 12, 0.02, 60
    K, mass = _chain(n)
    x0 = np.zeros(3 * n)
    x0[0::3] = np.arange(n, dtype=float)
    u0 = 0.05 * np.sin(np.linspace(0, np.pi, n))

    def fint(U):
        u = U.numpy()[0::3] - x0[0::3]
        out = np.zeros(3 * n)
        out[0::3] = K @ u
        return T(out)

    st = NewmarkJFNK(fint, T(x0), T(mass), dt, rtol=1e-12)
    st.x[0::3] += T(u0)
    st.xn = st.x.clone()
    a0 = -(K @ u0) / mass                         # a consistent initial acceleration
    st.xdd[0::3] = T(a0)
    # direct Newmark (beta = 1/4, gamma = 1/2) on M u'' + K u = 0 from the same state
    Mm = np.diag(mass)
    u, v, a = u0.copy(), np.zeros(n), a0.copy()
    beta, gam = 0.25, 0.5
    Keff = K + Mm / (beta * dt * dt)
    E0 = 0.5 * u @ K @ u
    for s in range(steps):
        info = st.step()
        assert info.converged, info
        rhs = Mm @ (u / (beta * dt * dt) + v / (beta * dt) + (1 / (2 * beta) - 1) * a)
        un = np.linalg.solve(Keff, rhs)
        an = (un - u) / (beta * dt * dt) - v / (beta * dt) - (1 / (2 * beta) - 1) * a
        vn = v + dt * ((1 - gam) * a + gam * an)
        u, v, a = un, vn, an
        got = st.x.numpy()[0::3] - x0[0::3]
        assert np.allclose(got, u, rtol=0, atol=1e-10), s
        assert np.allclose(st.xd.numpy()[0::3], v, rtol=0, atol=1e-8), s
    E = 0.5 * v @ Mm @ v + 0.5 * u @ K @ u
    vs = st.xd.numpy()[0::3]
    Es = 0.5 * vs @ Mm @ vs + 0.5 * got @ K @ got
    assert abs(E - E0) <= 1e-10 * E0 and abs(Es - E0) <= 1e-7 * E0


def test_damping_and_fixed_dofs():
    n, dt = 12, 0.02
    K, mass = _chain(n)
    x0 = np.zeros(3 * n)

    def fint(U):
        out = np.zeros(3 * n)
        out[0::3] = K @ U.numpy()[0::3]
        return T(out)

    fixed = np.zeros(3 * n, bool)
    fixed[0] = True                               # node 0 held in x
    st = NewmarkJFNK(fint, T(x0), T(mass), dt, fixed=torch.as_tensor(fixed), damping=True,
                     dampfactor=2.0, rtol=1e-12)
    st.x[3::3] = 0.02
    st.x[0] = 0.01                                # off its reference value: pulled back
    st.xn = st.x.clone()
    energies = []
    for _ in range(40):
        assert st.step().converged
        u, v = st.x.numpy()[0::3], st.xd.numpy()[0::3]
        energies.append(0.5 * v @ (mass * v) + 0.5 * u @ K @ u)
        assert abs(st.x[0].item()) <= 1e-10       # held by its residual row, to the tolerance
    assert all(e2 <= e1 * (1 + 1e-12) for e1, e2 in zip(energies[1:], energies[2:]))


def test_dampfactor_needs_damping_flag():
    """-dampfactor alone adds no damping: FDynamic applies Damp only under -damping
    (src/main.c:1257, :1326), so the residual is the undamped one, step for step."""
    n, dt = 8, 0.02
    K, mass = _chain(n)

    def fint(U):
        out = np.zeros(3 * n)
        out[0::3] = K @ U.numpy()[0::3]
        return T(out)

    runs = {}
    for key, kw in (("plain", {}), ("factor_only", dict(dampfactor=3.0)),
                    ("damped", dict(damping=True, dampfactor=3.0))):
        st = NewmarkJFNK(fint, T(np.zeros(3 * n)), T(mass), dt, rtol=1e-12, **kw)
        st.x[3::3] = 0.02
        st.xn = st.x.clone()
        for _ in range(5):
            assert st.step().converged
        runs[key] = st.x.numpy().copy()
        assert torch.count_nonzero(st.C).item() == (n * 3 if key == "damped" else 0)
    assert np.array_equal(runs["plain"], runs["factor_only"])
    assert not np.allclose(runs["plain"], runs["damped"], rtol=0, atol=1e-9)


def test_thickness_hook_runs_once_per_converged_step():
    """UpdateElementThickness is called after every converged step, never per Newton
    iteration (src/main.c:497-503); the hook's value is kept on the stepper."""
    n, dt = 6, 0.02
    K, mass = _chain(n)
    last = {}

    def fint(U):
        last["x"] = U.numpy()[0::3].copy()
        out = np.zeros(3 * n)
        out[0::3] = K @ U.numpy()[0::3]
        return T(out)

    calls = []

    def thickness():
        calls.append(last["x"].copy())
        return np.full(n - 1, float(len(calls)))

    st = NewmarkJFNK(fint, T(np.zeros(3 * n)), T(mass), dt, element_thickness=thickness,
                     rtol=1e-12)
    st.x[3::3] = 0.02
    st.xn = st.x.clone()
    for k in range(3):
        assert st.step().converged
        assert len(calls) == k + 1 and st.thickness[0] == k + 1
        assert np.array_equal(calls[-1], st.x.numpy()[0::3])   # evaluated at the accepted state

