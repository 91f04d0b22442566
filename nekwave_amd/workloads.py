"""The benchmark workloads (seeded synthetic inputs), as product-side builders.

    c1  PEC unit cube [0,1]^3, 4^3 hexes, p = 4 (n1 = 5), TM110 Ez = sin(pi x) sin(pi y),
        upwind, CFL 0.5 (dt = 0.5 dxmin), 200 steps
    c2  periodic [0,1]^3, 16^3 hexes, interior vertices jittered +-10% h (seed 2),
        p = 7 (n1 = 8), Ez = exp(-|x - 0.5|^2 / 0.1^2), upwind, automatic dt, 1000 steps
    c3  PEC [0,1]^3, 32^3 hexes, p = 7, eps_e ~ U[1,4] in element order (seed 3), mu = 1,
        TM110, automatic dt (c_max = 1), 500 steps
    c4  256 PEC unit-cube cavities; per instance, in draw order (seed 4): k ~ U{2..12}
        (k^3 hexes), p ~ U{3..8}, eps ~ U[1,4], m ~ U{1..3}, n ~ U{1..3}, phase ~ U[0, 2pi);
        TM_mn0 field at t = 0; common dt = the smallest automatic dt; 500 steps
    c5  [0,1]^3, 64^3 hexes, p = 5 (n1 = 6); seed 5: 6 uniforms place two Ez pulses of
        width 0.1 at U[0.2, 0.8]^3, then per element (element order) 4 uniforms give
        eps = exp(0.3 N), mu = exp(0.3 N) with N = Box-Muller(1 - u_a, u_b); PEC on y and z;
        the x faces are cut faces fed by the on-device pack -> copy -> unpack; upwind,
        automatic dt at the fastest wave speed max 1/sqrt(eps mu), 200 steps

Random draws: std::mt19937_64(seed), uniform = (x >> 11) 2^-53, U{a..b} = a +
floor(u (b - a + 1)), normal = Box-Muller first variate (nw_rand_*).
"""
import ctypes as C
import math

import numpy as np

from . import Mesh
from ._lib import check, lib


def uniforms(seed, n):
    out = np.zeros(n)
    check(lib().nw_rand_uniform(int(seed), int(n), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def normals(seed, n):
    out = np.zeros(n)
    check(lib().nw_rand_normal(int(seed), int(n), out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def tm_mode(x, y, t=0.0, m=1, n=1, eps=1.0, mu=1.0, phase=0.0):
    """TM_mn0 mode of the PEC unit cube: Ez = sin(m pi x) sin(n pi y) cos(w t + phase), w = pi sqrt(m^2+n^2)/sqrt(eps mu)."""
    om = math.pi * math.sqrt(m * m + n * n) / math.sqrt(eps * mu)
    c, s = math.cos(om * t + phase), math.sin(om * t + phase)
    sx, cx = np.sin(m * math.pi * x), np.cos(m * math.pi * x)
    sy, cy = np.sin(n * math.pi * y), np.cos(n * math.pi * y)
    z = np.zeros_like(x)
    return [z, z, sx * sy * c, -(n * math.pi / (mu * om)) * sx * cy * s, (m * math.pi / (mu * om)) * cx * sy * s, z]


def pack(fields):
    return np.concatenate([np.asarray(f, dtype=np.float64).ravel() for f in fields])


class Workload:
    def __init__(self, name, meshes, states, dt, steps, c0=1.0, precision="fp32", info=None):
        self.name = name
        self.meshes = meshes
        self.states = states
        self.dt = dt
        self.steps = steps
        self.c0 = c0
        self.precision = precision
        self.info = info or {}

    @property
    def npts(self):
        return sum(m.getTotalPoints() for m in self.meshes)

    @property
    def nelt(self):
        return sum(m.getNumElements() for m in self.meshes)


def c1():
    m = Mesh(5)
    m.createBoxMesh(4, 4, 4, 0, 1, 0, 1, 0, 1)
    x, y, _ = m.coords()
    return Workload("c1", [m], [pack(tm_mode(x, y))], 0.5 * m.computeMinNodeDistance(), 200, precision="fp64",
                    info={"description": "PEC unit cube 4^3, p=4, TM110, CFL 0.5, upwind"})


def c2():
    m = Mesh(8)
    m.createBoxMesh(16, 16, 16, 0, 1, 0, 1, 0, 1, True, True, True, jitter=0.1, seed=2)
    x, y, z = m.coords()
    r2 = (x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2
    zero = np.zeros_like(x)
    st = pack([zero, zero, np.exp(-r2 / 0.01), zero, zero, zero])
    return Workload("c2", [m], [st], m.computeAutomaticDt(1.0, 0.4), 1000,
                    info={"description": "periodic 16^3 jittered +-10% h (seed 2), p=7, Gaussian pulse, upwind"})


def c3():
    m = Mesh(8)
    m.createBoxMesh(32, 32, 32, 0, 1, 0, 1, 0, 1)
    eps = 1.0 + 3.0 * uniforms(3, 32 ** 3)
    m.setMaterials(eps, np.ones(32 ** 3))
    x, y, _ = m.coords()
    return Workload("c3", [m], [pack(tm_mode(x, y))], m.computeAutomaticDt(1.0, 0.4), 500,
                    info={"description": "PEC 32^3, p=7, eps_e ~ U[1,4] (seed 3), TM110, upwind"})


def c4(count=256):
    u = uniforms(4, 6 * count)
    meshes, states, dts, spec = [], [], [], []
    for x in range(count):
        d = u[6 * x:6 * x + 6]
        k = 2 + int(d[0] * 11)
        p = 3 + int(d[1] * 6)
        eps = 1.0 + 3.0 * d[2]
        mm = 1 + int(d[3] * 3)
        nn = 1 + int(d[4] * 3)
        ph = 2.0 * math.pi * d[5]
        m = Mesh(p + 1)
        m.createBoxMesh(k, k, k, 0, 1, 0, 1, 0, 1)
        m.setMaterials(np.full(k ** 3, eps), np.ones(k ** 3))
        xs, ys, _ = m.coords()
        meshes.append(m)
        states.append(pack(tm_mode(xs, ys, 0.0, mm, nn, eps, 1.0, ph)))
        dts.append(m.computeAutomaticDt(1.0, 0.4))
        spec.append((k, p, eps, mm, nn, ph))
    return Workload("c4", meshes, states, min(dts), 500,
                    info={"description": "%d PEC cavities, k~U{2..12}, p~U{3..8}, eps~U[1,4], TM_mn0 (seed 4)" % count,
                          "instances": spec})


def c5(nel=64):
    m = Mesh(6)
    m.createBoxMesh(nel, nel, nel, 0, 1, 0, 1, 0, 1, True, False, False)
    nelt = nel ** 3
    u = uniforms(5, 6 + 4 * nelt)
    cen = 0.2 + 0.6 * u[:6]
    w = u[6:].reshape(nelt, 4)

    def box_muller(a, b):
        return np.sqrt(-2.0 * np.log(1.0 - a)) * np.cos(2.0 * math.pi * b)

    eps = np.exp(0.3 * box_muller(w[:, 0], w[:, 1]))
    mu = np.exp(0.3 * box_muller(w[:, 2], w[:, 3]))
    m.setMaterials(eps, mu)
    m.selfHaloX()
    x, y, zc = m.coords()
    ez = np.zeros_like(x)
    for c in (cen[0:3], cen[3:6]):
        ez += np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (zc - c[2]) ** 2) / 0.01)
    zero = np.zeros_like(x)
    st = pack([zero, zero, ez, zero, zero, zero])
    # the automatic dt rule at the fastest wave speed of the medium, 1/sqrt(eps mu)
    # (eps mu reaches ~0.15 among 64^3 lognormal draws: dt at speed 1 is unstable)
    c_max = float(np.max(1.0 / np.sqrt(eps * mu)))
    return Workload("c5", [m], [st], m.computeAutomaticDt(c_max, 0.4), 200,
                    info={"description": "64^3 share, p=5, lognormal(0,0.3) eps and mu (seed 5), two pulses, "
                                         "PEC y/z, x cut faces through the on-device halo stand-in"})


BUILDERS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5}
