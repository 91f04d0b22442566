"""Closed-form Maxwell solutions and discrete norms used as test oracles.

All formulas are derived from the source-free Maxwell curl equations
eps dE/dt = curl H, mu dH/dt = -curl E; the reference's analytic tests
(tests/3dboxpec_test/3dboxpec_test.cpp:71-119, tests/3dboxper_test/3dboxper_test.cpp:57-103)
are restated here on box meshes because their .rea meshes are not in the tree.
"""
import math

import numpy as np


def tm_mode(x, y, t, m=1, n=1, eps=1.0, mu=1.0, phase=0.0):
    """TM_mn0 mode of the PEC unit cube [0,1]^3.  Returns (Ex,Ey,Ez,Hx,Hy,Hz)."""
    om = math.pi * math.sqrt(m * m + n * n) / math.sqrt(eps * mu)
    c, s = math.cos(om * t + phase), math.sin(om * t + phase)
    sx, cx = np.sin(m * math.pi * x), np.cos(m * math.pi * x)
    sy, cy = np.sin(n * math.pi * y), np.cos(n * math.pi * y)
    z = np.zeros_like(x)
    Ez = sx * sy * c
    Hx = -(n * math.pi / (mu * om)) * sx * cy * s
    Hy = +(m * math.pi / (mu * om)) * cx * sy * s
    return z, z.copy(), Ez, Hx, Hy, z.copy()


def boxpec_mode(x, y, z, t):
    """The (1,1,1) mode on [-1,1]^3 of 3dboxpec_test.cpp:29-47 (IC) and :71-92 (exact)."""
    om = math.pi * math.sqrt(3.0)
    th = math.cos(om * t) / math.sqrt(6.0)
    te = math.sin(om * t) / math.sqrt(2.0)
    pi = math.pi
    Ex = -np.cos(pi * x) * np.sin(pi * y) * np.sin(pi * z) * te
    Ey = np.sin(pi * x) * np.cos(pi * y) * np.sin(pi * z) * te
    Ez = np.zeros_like(x)
    Hx = -np.sin(pi * x) * np.cos(pi * y) * np.cos(pi * z) * th
    Hy = -np.cos(pi * x) * np.sin(pi * y) * np.cos(pi * z) * th
    Hz = 2.0 * np.cos(pi * x) * np.cos(pi * y) * np.sin(pi * z) * th
    return Ex, Ey, Ez, Hx, Hy, Hz


def boxper_mode(x, y, z, t):
    """The omega = sqrt(3) periodic mode on [0,2pi]^3 of 3dboxper_test.cpp:25-35, :57-77."""
    om = math.sqrt(3.0)
    th = math.sin(om * t) / om
    te = math.cos(om * t)
    Ex = np.zeros_like(x)
    Ey = np.cos(x) * np.sin(y) * np.sin(z) * te
    Ez = np.cos(x) * np.cos(y) * np.cos(z) * te
    Hx = 2.0 * np.cos(x) * np.sin(y) * np.cos(z) * th
    Hy = -np.sin(x) * np.cos(y) * np.cos(z) * th
    Hz = np.sin(x) * np.sin(y) * np.sin(z) * th
    return Ex, Ey, Ez, Hx, Hy, Hz


def pack(fields):
    return np.concatenate([np.asarray(f, dtype=np.float64).ravel() for f in fields])


def rel_l2(mesh_w3_jac, state, exact):
    """Relative L2 with dV = J*w3 (3dboxpec_test.cpp:76-119)."""
    dV = mesh_w3_jac
    n = dV.size
    num = 0.0
    den = 0.0
    for c in range(6):
        d = state[c * n:(c + 1) * n] - exact[c * n:(c + 1) * n]
        num += np.sum(d * d * dV)
        den += np.sum(exact[c * n:(c + 1) * n] ** 2 * dV)
    return math.sqrt(num / (den if den > 1e-14 else 1.0))


def dV(mesh):
    """J * w3 per node for an oracle mesh."""
    return mesh.jac * np.tile(mesh.w3, mesh.nelt)


def lsrk_amplification(a, b, z):
    """Apply one LSRK45 step (k = a k + dt f; u += b k) to u' = lambda u; z = lambda dt."""
    u = 1.0 + 0j
    k = 0.0 + 0j
    for s in range(5):
        k = a[s] * k + z * u
        u = u + b[s] * k
    return u
