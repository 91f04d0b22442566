"""Writes the small golden fixtures under tests/golden/ from the FP64 oracle.

    python tests/golden/make_golden.py

Each fixture holds the inputs (mesh parameters, seeded state) and the outputs
of the oracle: per-kernel results on that state, the state after one stage and
after one full LSRK45 step.  They freeze the oracle (a regression pin; the
oracle itself is pinned against the reference by tests/test_oracle_*.py) and
give the GPU tests a fixed target that needs no oracle build on the box.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

CASES = {
    # name: n1, nel, bounds, periodic, jitter, seed, media, c0
    "one_element_n3": (3, (1, 1, 1), (0, 1, 0, 1, 0, 1), (0, 0, 0), 0.0, 0, False, 1.0),
    "box2_n4_pec": (4, (2, 2, 2), (-0.5, 0.5, -0.5, 0.5, -0.5, 0.5), (0, 0, 0), 0.0, 0, False, 1.0),
    "box2_n5_periodic_jitter": (5, (2, 2, 2), (0, 1, 0, 1, 0, 1), (1, 1, 1), 0.1, 2, False, 1.0),
    "box2_n4_media_central": (4, (2, 2, 2), (0, 1, 0, 1, 0, 1), (0, 1, 0), 0.05, 9, True, 0.0),
}


def build(name):
    n1, nel, b, per, jit, seed, media, c0 = CASES[name]
    o = po.Mesh.box(n1, nel, b, per, jitter=jit, seed=seed)
    eps = mu = None
    if media:
        r = po.Rng(17)
        eps = np.array([1.0 + 3.0 * r.uniform() for _ in range(o.nelt)])
        mu = np.array([0.5 + 1.5 * r.uniform() for _ in range(o.nelt)])
        o.set_materials(eps, mu)
    return o, eps, mu, c0


def seeded_state(o, seed=123):
    r = po.Rng(seed)
    return np.array([r.uniform() * 2.0 - 1.0 for _ in range(6 * o.npts)])


def main():
    for name in CASES:
        o, eps, mu, c0 = build(name)
        st = seeded_state(o)
        n = o.npts
        vol_e = np.concatenate(po.volume_curl(o, st[3 * n:4 * n], st[4 * n:5 * n], st[5 * n:6 * n], 1.0))
        fEN, fHN = po.restrict(o, st)
        fl = po.flux(o, st, fEN, fHN, c0)
        dt = 0.5 * o.auto_dt()
        s = po.Solver(o, c0)
        s.upload(st)
        s.stage(0, dt)
        after_stage = s.download()
        s = po.Solver(o, c0)
        s.upload(st)
        s.step(dt)
        after_step = s.download()
        out = dict(n1=o.n1, nel=np.array(CASES[name][1]), bounds=np.array(CASES[name][2], dtype=float),
                   periodic=np.array(CASES[name][3]), jitter=CASES[name][4], seed=CASES[name][5], c0=c0,
                   eps=eps if eps is not None else np.zeros(0), mu=mu if mu is not None else np.zeros(0),
                   state=st, dt=dt, volume_curl_H=vol_e, flux=fl, after_stage=after_stage, after_step=after_step,
                   energy=o.energy(st))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
        print(name, os.path.getsize(os.path.join(HERE, name + ".npz")))


if __name__ == "__main__":
    main()
