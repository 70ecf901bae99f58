"""The synthetic workloads the force path is measured and checked on.

Every generator is deterministic (numpy default_rng seeds as stated).  Material mapping
for the reference's Neo-Hookean active-strain law: mu = E / (2 (1 + nu)),
K = E / (3 (1 - 2 nu)); "3-point quadrature" is n_qp = 3 through-thickness Gauss points
(the reference has one in-plane point per element, src/active_strain.c:69-88, 196-201).
"""
from __future__ import annotations

import numpy as np

from .fem import ActParams
from .mesh import Mesh, centroids, flat_membrane


def membrane_small(seed: int = 0):
    """Flat 1.0 x 0.75 membrane, nx=32 x ny=24 (1536 triangles), jitter +-1e-3 h (seed 0),
    E=1, nu=0.3, h0=0.01, uniform active strain 0.1 along x (cart_fib_act, f = e_x)."""
    m = flat_membrane(32, 24, 1.0, 0.75, jitter=1e-3, seed=seed)
    p = ActParams.from_young(1.0, 0.3, n_qp=3, c33_iters=3, cart_fib_act=1, h0=0.01,
                             gamma=0.1, ramp_type=0)
    x, y, z = m.x0.copy(), m.y0.copy(), m.z0.copy()
    return m, p, (x, y, z)


def membrane_large(nx: int = 256, ny: int = 192, seed: int = 1):
    """Flat 1.0 x 0.75 membrane nx x ny (98 304 triangles at 256 x 192).  gamma_e =
    0.15 sin(pi x_c) cos(pi y_c / 2) at the reference centroid; fibre angle U[-30, 30] deg
    per element; current z perturbed by N(0, 1e-3) per real node.  default_rng(seed) draws
    the fibre angles first, then the z perturbation."""
    m = flat_membrane(nx, ny, 1.0, 0.75, jitter=1e-3, seed=0)
    rng = np.random.default_rng(seed)
    ang = np.deg2rad(rng.uniform(-30.0, 30.0, m.n_elmt))
    m.n_fib = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], 1)
    c = centroids(m)
    m.gamma_scale = np.sin(np.pi * c[:, 0]) * np.cos(np.pi * c[:, 1] / 2.0)
    p = ActParams.from_young(1.0, 0.3, n_qp=3, c33_iters=3, cart_fib_act=1, h0=0.01,
                             gamma=0.15, ramp_type=0)
    x, y, z = m.x0.copy(), m.y0.copy(), m.z0.copy()
    z[: m.n_v] += rng.normal(0.0, 1e-3, m.n_v)
    m.apply_ghosts(x, y, z)
    return m, p, (x, y, z)


def smooth_displacement(m: Mesh, amp: float = 0.01):
    """Current coordinates X + a smooth in-plane + out-of-plane field (ghosts by rule)."""
    x, y, z = m.x0.copy(), m.y0.copy(), m.z0.copy()
    X, Y = m.x0[: m.n_v], m.y0[: m.n_v]
    x[: m.n_v] += amp * np.sin(np.pi * X) * np.cos(np.pi * Y)
    y[: m.n_v] += 0.5 * amp * np.sin(2 * np.pi * Y) * X
    z[: m.n_v] += amp * np.sin(np.pi * X) * np.sin(np.pi * Y / 0.75)
    m.apply_ghosts(x, y, z)
    return x, y, z


def ellipsoid_shell(levels: int = 6, ti: int = 30, seed: int = 3):
    """Truncated prolate ellipsoid (semi-axes 1, 1, 2.2; z >= 0.5), 184 320 triangles at
    levels = 6, valence-5/7 extraordinary vertices (irregular Loop patches), helix -60 -> +60
    deg apex to rim, active strain 0.12 times a Gaussian pulse in time (ramp type 4,
    t0 = 0.3, sigma = 0.1, dt = 0.01).  Current state: the reference shell contracted by 2 %
    towards the axis and 1 % along it, so the force is not the reference-state force."""
    from .mesh import ellipsoid_dome
    m = ellipsoid_dome(levels=levels, seed=seed)
    p = ActParams.from_young(1.0, 0.3, n_qp=3, c33_iters=3, cart_fib_act=1, h0=0.01,
                             gamma=0.12, ramp_type=4, pulse_t0=0.3, pulse_sigma=0.1, ti=ti,
                             dt=0.01, tisteps=100)
    x, y, z = m.x0.copy(), m.y0.copy(), m.z0.copy()
    n = m.n_v
    x[:n] *= 0.98
    y[:n] *= 0.98
    z[:n] *= 0.99
    m.apply_ghosts(x, y, z)
    return m, p, (x, y, z)


def batch_membranes(n: int = 64, seed: int = 2, nx_lo: int = 16, nx_hi: int = 256,
                    amp: float = 0.01):
    """n independent flat 1.0 x 0.75 membranes for one batched launch: nx log-uniform in
    [nx_lo, nx_hi], ny = round(0.75 nx); per membrane Young's modulus U[0.5, 2] (nu = 0.3),
    thickness U[0.005, 0.02], active strain U[0.05, 0.2] along a fibre at U[-30, 30] deg per
    element; current state = reference + the smooth displacement field of amplitude amp.
    default_rng(seed) draws nx, E, h0, gamma for all membranes first, then per membrane its
    fibre angles; mesh jitter uses seed 1000 + i."""
    rng = np.random.default_rng(seed)
    nxs = np.rint(np.exp(rng.uniform(np.log(nx_lo), np.log(nx_hi), n))).astype(int)
    E_mod = rng.uniform(0.5, 2.0, n)
    h0 = rng.uniform(0.005, 0.02, n)
    gam = rng.uniform(0.05, 0.2, n)
    out = []
    for i in range(n):
        nx = int(nxs[i])
        ny = max(2, int(round(0.75 * nx)))
        m = flat_membrane(nx, ny, 1.0, 0.75, jitter=1e-3, seed=1000 + i)
        ang = np.deg2rad(rng.uniform(-30.0, 30.0, m.n_elmt))
        m.n_fib = np.stack([np.cos(ang), np.sin(ang), np.zeros_like(ang)], 1)
        p = ActParams.from_young(float(E_mod[i]), 0.3, n_qp=3, c33_iters=3, cart_fib_act=1,
                                 h0=float(h0[i]), gamma=float(gam[i]), ramp_type=0)
        out.append((m, p, smooth_displacement(m, amp)))
    return out
