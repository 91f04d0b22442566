"""Pins the oracle's diagnostics (row A16: case.cpp:733-794 with the eps/mu
weighting of the media extension) against closed forms and a brute force.

  * max |E|, max |H|: a seeded state whose E and H components have different
    scales and a planted maximum in each component, against numpy's reduction
    over the node magnitudes; a component-offset slip (H read from Ez, Hx, Hy) or
    a max over components instead of magnitudes fails it.  TM110 at t = 0 has
    max |E| = 1 at the node (1/2, 1/2, z); at a quarter period max |H| = 1/sqrt(2).
  * energy with media: constant fields on affine elements integrate exactly, so
    U = 1/2 sum_e (eps_e |E|^2 + mu_e |H|^2) vol_e, also with a field that
    differs between elements; swapping eps and mu, dropping the 1/2 or weighting
    by the wrong element fails.  A homogeneous-eps TM_mn0
    mode has U = eps/8 both at t = 0 (all electric) and at a quarter period (all
    magnetic), which pins the eps and mu weights through the physics.
"""
import math

import numpy as np
import pytest

import analytic as an
import pyoracle as po

UNIT = (0, 1, 0, 1, 0, 1)


def test_max_e_h_against_brute_force():
    m = po.Mesh.box(4, (2, 2, 2), UNIT, jitter=0.1, seed=7)
    n = m.npts
    rng = np.random.default_rng(11)
    u = np.concatenate([rng.uniform(-1.0, 1.0, n) * s for s in (1.0, 2.0, 0.5, 3.0, 0.25, 4.0)])
    # a planted maximum in each component, each at its own node
    for c, (node, v) in enumerate([(3, 9.0), (17, -8.5), (29, 8.0), (41, 20.0), (53, -19.0), (65, 18.5)]):
        u[c * n + node] = v
    comps = u.reshape(6, n)
    want_e = np.max(np.sqrt(comps[0] ** 2 + comps[1] ** 2 + comps[2] ** 2))
    want_h = np.max(np.sqrt(comps[3] ** 2 + comps[4] ** 2 + comps[5] ** 2))
    assert m.max_e(u) == pytest.approx(want_e, rel=1e-15)
    assert m.max_h(u) == pytest.approx(want_h, rel=1e-15)
    assert abs(want_e - want_h) > 5.0  # an E/H swap or offset would not pass


def test_max_e_h_tm110_closed_form():
    m = po.Mesh.box(5, (4, 4, 4), UNIT)  # element boundaries at x, y = 1/2 carry nodes
    x, y = m.x, m.y
    u0 = an.pack(an.tm_mode(x, y, 0.0))
    assert m.max_e(u0) == 1.0  # sin(pi/2)^2 at (1/2, 1/2)
    assert m.max_h(u0) == 0.0
    om = math.pi * math.sqrt(2.0)
    uq = an.pack(an.tm_mode(x, y, 0.5 * math.pi / om))  # quarter period: all magnetic
    assert m.max_h(uq) == pytest.approx(1.0 / math.sqrt(2.0), rel=1e-14)  # at (1/2, 0) and (0, 1/2)
    assert m.max_e(uq) < 1e-15


def test_energy_constant_fields_with_media_exact():
    nel = (3, 2, 2)
    m = po.Mesh.box(4, nel, (0.0, 1.5, 0.0, 1.0, 0.0, 2.0))
    nelt = m.nelt
    rng = np.random.default_rng(5)
    eps = rng.uniform(1.0, 4.0, nelt)
    mu = rng.uniform(0.5, 3.0, nelt)
    m.set_materials(eps, mu)
    n = m.npts
    one = np.ones(n)
    e2, h2 = 1.0 + 4.0 + 0.25, 9.0 + 0.0 + 1.0  # |(1, 2, 0.5)|^2 and |(3, 0, 1)|^2
    u = np.concatenate([one, 2 * one, 0.5 * one, 3 * one, 0 * one, one])
    vol = (1.5 / 3) * (1.0 / 2) * (2.0 / 2)
    want = 0.5 * vol * float(np.sum(eps * e2 + mu * h2))
    assert m.energy(u) == pytest.approx(want, rel=1e-13)
    # E only and H only separate the two weights
    ue = np.concatenate([one, 2 * one, 0.5 * one, 0 * one, 0 * one, 0 * one])
    uh = np.concatenate([0 * one, 0 * one, 0 * one, 3 * one, 0 * one, one])
    assert m.energy(ue) == pytest.approx(0.5 * vol * e2 * float(np.sum(eps)), rel=1e-13)
    assert m.energy(uh) == pytest.approx(0.5 * vol * h2 * float(np.sum(mu)), rel=1e-13)
    # a field constant per element but different between elements: each element's
    # eps and mu must weigh its own field
    a = np.repeat(np.arange(1.0, nelt + 1.0), m.np)
    ua = np.concatenate([a, 0 * one, 0 * one, 0 * one, 0.5 * a, 0 * one])
    want = 0.5 * vol * float(np.sum(eps * np.arange(1.0, nelt + 1.0) ** 2 + mu * 0.25 * np.arange(1.0, nelt + 1.0) ** 2))
    assert m.energy(ua) == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("eps,mu,mn", [(2.0, 3.0, (1, 1)), (3.5, 0.5, (2, 1)), (1.0, 1.0, (1, 3))])
def test_energy_tm_mode_with_media_is_eps_over_8(eps, mu, mn):
    m = po.Mesh.box(8, (4, 4, 2), UNIT)
    m.set_materials(np.full(m.nelt, eps), np.full(m.nelt, mu))
    x, y = m.x, m.y
    mm, nn = mn
    om = math.pi * math.sqrt(mm * mm + nn * nn) / math.sqrt(eps * mu)
    u0 = an.pack(an.tm_mode(x, y, 0.0, mm, nn, eps, mu))
    uq = an.pack(an.tm_mode(x, y, 0.5 * math.pi / om, mm, nn, eps, mu))
    assert m.energy(u0) == pytest.approx(eps / 8.0, rel=1e-9)  # (eps/2) * 1/4
    assert m.energy(uq) == pytest.approx(eps / 8.0, rel=1e-9)  # (mu/2) * eps / (4 mu)
