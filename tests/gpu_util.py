"""Shared helpers of the GPU tests: build the same case for the product (HIP,
through the C ABI) and for the oracle (CPU), run both, compare."""
import math

import numpy as np

import analytic as an
import nekwave_amd as nw
import pyoracle as po


def pair(n1, nel, bounds=(0, 1, 0, 1, 0, 1), per=(0, 0, 0), jitter=0.0, seed=0, eps=None, mu=None, halo=False):
    m = nw.Mesh(n1)
    m.createBoxMesh(*nel, *bounds, *[bool(p) for p in per], jitter=jitter, seed=seed)
    o = po.Mesh.box(n1, nel, bounds, per, jitter=jitter, seed=seed)
    if eps is not None or mu is not None:
        m.setMaterials(eps, mu)
        o.set_materials(eps, mu)
    if halo:
        m.selfHaloX()
        o.self_halo_x()
    return m, o


def run_product(m, st0, dt, nsteps, c0=1.0, precision="fp64", pipeline="fused", use_graph=True, tile_option=-1):
    s = nw.DgSolver(precision=precision, pipeline=pipeline, use_graph=use_graph, tile_option=tile_option)
    s.initialize(m, c0)
    s.uploadState(st0)
    if nsteps:
        s.run(dt, nsteps)
    out = s.downloadState()
    s.finalize()
    return out


def run_oracle(o, st0, dt, nsteps, c0=1.0):
    s = po.Solver(o, c0)
    s.upload(st0)
    s.step(dt, nsteps)
    return s.download()


def rel_l2(o, a, b):
    dV = an.dV(o)
    n = o.npts
    num = den = 0.0
    for c in range(6):
        d = a[c * n:(c + 1) * n] - b[c * n:(c + 1) * n]
        num += np.sum(d * d * dV)
        den += np.sum(b[c * n:(c + 1) * n] ** 2 * dV)
    return math.sqrt(num / den)


def gaussian_state(o, centre=(0.5, 0.5, 0.5), width=0.1):
    x, y, z = o.x, o.y, o.zc
    r2 = (x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2
    zero = np.zeros_like(x)
    return an.pack([zero, zero, np.exp(-r2 / width ** 2), zero, zero, zero])


def random_state(o, seed=0):
    rng = np.random.default_rng(seed)
    x, y, z = o.x, o.y, o.zc
    f = [np.sin(2 * math.pi * (x + rng.uniform())) * np.cos(math.pi * y) for _ in range(6)]
    return an.pack(f) + 0.1 * rng.standard_normal(6 * o.npts)


def lift_numpy(vM, dA, fl, rhs0, T, npts):
    """rhs[c npts + vM[p]] += dA[p] F_c[p] with the E rows taken from the flux's
    second half (dg_kernels.cu:316-338), accumulated point by point with np.add.at."""
    out = rhs0.copy()
    for c, src in enumerate((3, 4, 5, 0, 1, 2)):
        np.add.at(out, c * npts + vM, dA * fl[src * T:(src + 1) * T])
    return out
