"""Pins the oracle's step (rows A7-A17) with what the mathematics fixes.

Each check is chosen so that a plausible slip (a transposed D, a swapped
metric term, a flipped flux sign, a wrong LSRK coefficient, a wrong face-node
map) fails it:
  * LSRK45: the stability polynomial of dg_solver.cpp:58-69 must be the
    4th-order Taylor polynomial + c5 z^5;
  * volume curl: exact for linear fields on jittered trilinear elements and for
    degree-p polynomials on affine elements;
  * flux/lift: zero jumps for continuous fields; the semi-discrete energy is
    conserved with the central flux and non-increasing with upwind
    (tests/nekwave_test.cpp:212-293);
  * full step: analytic eigenmodes (TM110 on the unit cube; the reference's
    3dboxpec and 3dboxper modes restated on box meshes, threshold 0.05 from
    3dboxpec_test.cpp:136), p-convergence, 4th-order time convergence;
  * extensions: media with eps = mu = 1 reproduce the reference bit for bit,
    homogeneous eps scales the frequency by 1/sqrt(eps), the self-halo run is
    bit-identical to the periodic run.
"""
import math

import numpy as np
import pytest

import analytic as an
import pyoracle as po

UNIT = (0, 1, 0, 1, 0, 1)


def _run(mesh, state0, dt, nsteps, c0=1.0):
    s = po.Solver(mesh, c0)
    s.upload(state0)
    s.step(dt, nsteps)
    return s.download()


# ------------------------------------------------------------------ LSRK ----

def test_lsrk45_stability_polynomial():
    a, b = po.rk_coeffs()
    # polynomial coefficients of R(z) through exact polynomial arithmetic
    u = np.polynomial.Polynomial([1.0])
    k = np.polynomial.Polynomial([0.0])
    zpoly = np.polynomial.Polynomial([0.0, 1.0])
    for s in range(5):
        k = a[s] * k + zpoly * u
        u = u + b[s] * k
    c = u.coef
    np.testing.assert_allclose(c[:5], [1, 1, 0.5, 1 / 6, 1 / 24], rtol=0, atol=2e-15)
    assert len(c) == 6 and 0 < c[5] < 1 / 120  # c5 of CK45 is about 1/200
    assert a[0] == 0.0


def test_lsrk_update_kernel_applies_the_recurrence():
    """or_lsrk_update (dg_kernels.cu:393-408) driven by rhs = lambda*u reproduces R(z)."""
    m = po.Mesh.box(3, (1, 1, 1), UNIT)
    a, b = po.rk_coeffs()
    lam, dt = -0.7, 0.3
    u = np.linspace(0.5, 1.5, 6 * m.npts)
    k = np.zeros_like(u)
    u0 = u.copy()
    for s in range(5):
        u, k = po.lsrk_update(m, u, k, lam * u, a[s], b[s], dt)
    np.testing.assert_allclose(u, an.lsrk_amplification(a, b, lam * dt).real * u0, rtol=1e-14)


# ---------------------------------------------------------- volume curl ----

def _weighted_to_curl(mesh, w):
    return w / an.dV(mesh)


def test_volume_curl_exact_for_linear_fields_on_trilinear_elements():
    m = po.Mesh.box(4, (3, 3, 3), UNIT, (0, 0, 0), jitter=0.1, seed=11)
    x, y, z = m.x, m.y, m.zc
    A = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [-0.6, 0.9, 1.5]])
    u = [A[c, 0] * x + A[c, 1] * y + A[c, 2] * z + c for c in range(3)]
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    for sign in (1.0, -1.0):
        w = po.volume_curl(m, *u, sign)
        for c in range(3):
            np.testing.assert_allclose(_weighted_to_curl(m, w[c]), sign * curl[c], atol=1e-11)


@pytest.mark.parametrize("n1", [3, 5, 8])
def test_volume_curl_exact_for_degree_p_polynomials_on_affine_elements(n1):
    p = n1 - 1
    m = po.Mesh.box(n1, (2, 1, 2), (-1, 0.5, 0, 2, 0.2, 1.1))
    x, y, z = m.x, m.y, m.zc
    u1 = y ** p + x * z ** (p - 1)
    u2 = z ** p - 2 * x ** (p - 1) * y
    u3 = x ** p * 0.5 + y * z
    # curl u = (d u3/dy - d u2/dz, d u1/dz - d u3/dx, d u2/dx - d u1/dy)
    c1 = z - p * z ** (p - 1)
    c2 = (p - 1) * x * z ** (p - 2) - 0.5 * p * x ** (p - 1)
    c3 = -2 * (p - 1) * x ** (p - 2) * y - p * y ** (p - 1)
    w = po.volume_curl(m, u1, u2, u3, 1.0)
    for got, ref in zip(w, (c1, c2, c3)):
        np.testing.assert_allclose(_weighted_to_curl(m, got), ref, atol=1e-9 * max(1.0, np.abs(ref).max()))


# --------------------------------------------------------- flux and lift ----

def test_continuous_field_has_zero_interior_flux():
    m = po.Mesh.box(4, (3, 2, 2), UNIT, (0, 0, 0), jitter=0.08, seed=5)
    x, y, z = m.x, m.y, m.zc
    st = an.pack([np.sin(x + y), x * z, np.cos(z), y * y, x - z, x * y * z])
    fEN, fHN = po.restrict(m, st)
    for c0 in (0.0, 1.0):
        fl = po.flux(m, st, fEN, fHN, c0).reshape(6, m.tfp)
        interior = m.vP >= 0
        # shared face nodes agree to rounding (the two sides sum the trilinear
        # map over their corners in different orders)
        assert np.max(np.abs(fl[:, interior])) < 1e-14
        assert np.max(np.abs(fl[:, ~interior])) > 1e-3


def test_pec_face_flux_kills_tangential_e():
    """On a PEC face the flux drives n x E to zero: F_E = -c0/2 n x (n x (-2 E-)) = -c0 (E- tangential)."""
    m = po.Mesh.box(3, (1, 1, 1), UNIT)
    rng = np.random.default_rng(0)
    st = rng.standard_normal(6 * m.npts)
    fEN, fHN = po.restrict(m, st)
    fl = po.flux(m, st, fEN, fHN, 1.0).reshape(6, m.tfp)
    E = fEN.reshape(3, m.tfp)
    n = np.stack([m.nx, m.ny, m.nz])
    Et = E - n * np.sum(n * E, axis=0)
    np.testing.assert_allclose(fl[3:6], -Et, atol=1e-14)
    nxE = np.cross(n.T, E.T).T
    np.testing.assert_allclose(fl[0:3], nxE, atol=1e-14)  # H-flux = -1/2 n x (-2E) = n x E


def _rhs(mesh, st, c0):
    """du/dt of the semi-discrete operator: stage 0 has a = 0, so k = dt * rhs."""
    s = po.Solver(mesh, c0)
    s.upload(st)
    s.stage(0, 1.0)
    return s.k()


def _energy_rate_from_faces(m, st, c0):
    """Exact semi-discrete energy balance of the strong-form DG operator:
    dU/dt = -c0/2 sum_interior dA (|[[E]]_t|^2/(Z+ + Z-) + |[[H]]_t|^2/(Y+ + Y-))
            - c0 sum_PEC dA |E-_t|^2 / Z-   (sums over every face point, both sides)."""
    S = st.reshape(6, m.npts)
    vM, vP = m.vM, m.vP
    n = np.stack([m.nx, m.ny, m.nz])
    eps = m.eps if m.eps.size else np.ones(m.nelt)
    mu = m.mu if m.mu.size else np.ones(m.nelt)
    Z = np.sqrt(mu / eps)
    tan = lambda v: v - n * np.sum(n * v, axis=0)
    Em, Hm = S[0:3, vM], S[3:6, vM]
    inner = vP >= 0
    Zm = Z[vM // m.np]
    Zp = np.where(inner, Z[np.maximum(vP, 0) // m.np], Zm)
    Ep, Hp = S[0:3, np.maximum(vP, 0)], S[3:6, np.maximum(vP, 0)]
    jE, jH = tan(Ep - Em), tan(Hp - Hm)
    inter = -0.5 * c0 * m.dA * (np.sum(jE * jE, 0) / (Zp + Zm) + np.sum(jH * jH, 0) / (1 / Zp + 1 / Zm))
    Et = tan(Em)
    pec = -c0 * m.dA * np.sum(Et * Et, 0) / Zm
    return np.sum(np.where(inner, inter, pec))


@pytest.mark.parametrize("c0", [0.0, 1.0])
@pytest.mark.parametrize("media", [False, True])
def test_semidiscrete_energy_balance(c0, media):
    """On affine elements GLL summation-by-parts makes the balance exact:
    central flux conserves energy, upwind dissipates exactly the jump energy.
    Pins flux signs and weights, the lift's face-node map and dA, inverse mass."""
    m = po.Mesh.box(5, (3, 2, 2), (0, 1.5, 0, 1, 0, 1), (0, 1, 0))
    rng = np.random.default_rng(3)
    if media:
        m.set_materials(rng.uniform(1, 4, m.nelt), rng.uniform(0.5, 2, m.nelt))
    st = rng.standard_normal(6 * m.npts)
    r = _rhs(m, st, c0).reshape(6, m.npts)
    S = st.reshape(6, m.npts)
    w = an.dV(m)
    eps = np.repeat(m.eps, m.np) if media else 1.0
    mu = np.repeat(m.mu, m.np) if media else 1.0
    dU = np.sum(w * (eps * np.sum(S[0:3] * r[0:3], 0) + mu * np.sum(S[3:6] * r[3:6], 0)))
    ref = _energy_rate_from_faces(m, st, c0)
    scale = np.sum(w * np.sum(np.abs(S) * np.abs(r), 0))
    assert abs(dU - ref) <= 1e-12 * scale, (dU, ref, scale)
    if c0 > 0:
        assert ref < -1e-3 * scale


def test_central_energy_drift_is_time_integration_error():
    """With the central flux on affine elements the energy drift over a fixed
    time falls with dt at high order (LSRK45 temporal error only; on trilinear
    elements the strong form is not exactly conservative, so the drift there
    has a spatial part)."""
    m = po.Mesh.box(4, (3, 3, 3), UNIT, (0, 0, 0))
    x, y, z = m.x, m.y, m.zc
    r2 = (x - 0.5) ** 2 + (y - 0.4) ** 2 + (z - 0.6) ** 2
    st = an.pack([0 * x, np.exp(-r2 / 0.05), 0 * x, 0 * x, 0 * x, np.exp(-r2 / 0.08)])
    e0 = m.energy(st)
    dt = m.auto_dt()
    d1 = abs(m.energy(_run(m, st, dt, 20, c0=0.0)) - e0) / e0
    d2 = abs(m.energy(_run(m, st, dt / 2, 40, c0=0.0)) - e0) / e0
    assert d1 < 1e-3 and d2 < d1 / 16, (d1, d2)  # about 2^5: LSRK45 dissipation
    assert m.energy(_run(m, st, dt, 20, c0=1.0)) < e0 * (1 - 1e-4)


def test_maxwell_stability_kat():
    """tests/nekwave_test.cpp:212-293 restated on the oracle."""
    N = 4
    m = po.Mesh.box(N, (2, 2, 2), (-0.5, 0.5, -0.5, 0.5, -0.5, 0.5))
    x, y, z = m.x, m.y, m.zc
    st = np.zeros(6 * m.npts)
    st[2 * m.npts:3 * m.npts] = np.exp(-50.0 * (x * x + y * y + z * z))
    e0 = m.energy(st)
    assert e0 > 0
    s = po.Solver(m, 1.0)
    s.upload(st)
    for _ in range(5):
        s.step(0.005)
        e = m.energy(s.download())
        assert e <= e0 + 1e-12 and math.isfinite(e)


# --------------------------------------------------------------- full step ----

def test_tm110_unit_cube_c1():
    """PEC unit cube 4^3, p=4, CFL 0.5 (dt = 0.0215841), upwind, 200 steps, TM110."""
    m = po.Mesh.box(5, (4, 4, 4), UNIT)
    dt = 0.5 * m.dxmin()
    assert abs(dt - 0.0215841) < 5e-8
    st0 = an.pack(an.tm_mode(m.x, m.y, 0.0))
    assert abs(m.energy(st0) - 0.125) < 1e-6
    out = _run(m, st0, dt, 200)
    ex = an.pack(an.tm_mode(m.x, m.y, 200 * dt))
    err = an.rel_l2(an.dV(m), out, ex)
    assert err < 2e-3, err
    drift = (m.energy(out) - m.energy(st0)) / m.energy(st0)
    assert -1e-3 < drift <= 0.0


def test_reference_boxpec_mode_restated():
    """3dboxpec_test: 3x3x3 on [-1,1]^3, order 4, central flux, auto dt, 50 steps, L2 <= 0.05."""
    m = po.Mesh.box(4, (3, 3, 3), (-1, 1, -1, 1, -1, 1))
    dt = m.auto_dt(1.0, 0.4)
    st0 = an.pack(an.boxpec_mode(m.x, m.y, m.zc, 0.0))
    out = _run(m, st0, dt, 50, c0=0.0)
    err = an.rel_l2(an.dV(m), out, an.pack(an.boxpec_mode(m.x, m.y, m.zc, 50 * dt)))
    assert err <= 0.05, err


def test_reference_boxper_mode_restated():
    """3dboxper_test: [0,2pi]^3 periodic 3x3x3, order 4, central flux, 50 steps, L2 <= 0.05."""
    L = 2 * math.pi
    m = po.Mesh.box(4, (3, 3, 3), (0, L, 0, L, 0, L), (1, 1, 1))
    dt = m.auto_dt(1.0, 0.4)
    st0 = an.pack(an.boxper_mode(m.x, m.y, m.zc, 0.0))
    out = _run(m, st0, dt, 50, c0=0.0)
    err = an.rel_l2(an.dV(m), out, an.pack(an.boxper_mode(m.x, m.y, m.zc, 50 * dt)))
    assert err <= 0.05, err


def test_p_convergence_tm110():
    T = 0.3
    errs = []
    for n1 in (3, 4, 5, 6):
        m = po.Mesh.box(n1, (2, 2, 2), UNIT)
        nst = int(math.ceil(T / (0.2 * m.dxmin())))
        dt = T / nst
        out = _run(m, an.pack(an.tm_mode(m.x, m.y, 0.0)), dt, nst)
        errs.append(an.rel_l2(an.dV(m), out, an.pack(an.tm_mode(m.x, m.y, T))))
    for e0, e1 in zip(errs, errs[1:]):
        assert e1 < e0 / 3, errs


def test_temporal_fourth_order():
    m = po.Mesh.box(5, (2, 2, 2), UNIT)
    st0 = an.pack(an.tm_mode(m.x, m.y, 0.0))
    T = 0.48
    ref = _run(m, st0, T / 96, 96)
    e = []
    for n in (6, 12):
        out = _run(m, st0, T / n, n)
        e.append(math.sqrt(np.sum((out - ref) ** 2)))
    rate = math.log2(e[0] / e[1])
    assert 3.6 < rate < 4.6, (e, rate)


# ----------------------------------------------------------- extensions ----

def test_unit_media_reproduce_reference_bitwise():
    m0 = po.Mesh.box(4, (2, 3, 2), UNIT, (1, 0, 1), jitter=0.1, seed=3)
    m1 = po.Mesh.box(4, (2, 3, 2), UNIT, (1, 0, 1), jitter=0.1, seed=3)
    m1.set_materials(np.ones(m1.nelt), np.ones(m1.nelt))
    x, y, z = m0.x, m0.y, m0.zc
    st = an.pack([np.sin(3 * x), y * z, np.exp(-x), np.cos(y), x * x, z])
    dt = m0.auto_dt()
    a = _run(m0, st, dt, 3)
    b = _run(m1, st, dt, 3)
    assert np.array_equal(a, b)
    assert m0.energy(a) == m1.energy(b)


def test_homogeneous_eps_scales_frequency():
    eps = 4.0
    m = po.Mesh.box(6, (3, 3, 3), UNIT)
    m.set_materials(np.full(m.nelt, eps), np.ones(m.nelt))
    st0 = an.pack(an.tm_mode(m.x, m.y, 0.0, eps=eps))
    dt = 0.5 * m.dxmin()
    n = 60
    out = _run(m, st0, dt, n)
    err = an.rel_l2(an.dV(m), out, an.pack(an.tm_mode(m.x, m.y, n * dt, eps=eps)))
    wrong = an.rel_l2(an.dV(m), out, an.pack(an.tm_mode(m.x, m.y, n * dt, eps=1.0)))
    assert err < 1e-4 and wrong > 100 * err, (err, wrong)


def test_impedance_flux_is_upwind_riemann_solution():
    """Two-medium interface: with the impedance-weighted upwind flux a right-going
    wave in medium 1 (H = Y1 n x E) meeting a left-going wave in medium 2 gives a
    flux that depends only on the incoming characteristics."""
    m = po.Mesh.box(3, (2, 1, 1), UNIT)
    eps = np.array([1.0, 3.0]); mu = np.array([1.0, 2.0])
    m.set_materials(eps, mu)
    rng = np.random.default_rng(1)
    st = rng.standard_normal(6 * m.npts)
    fEN, fHN = po.restrict(m, st)
    fl = po.flux(m, st, fEN, fHN, 1.0).reshape(6, m.tfp)
    # pick the interior face 1 of element 0 (normal +x), partner face 3 of element 1
    n2 = 9
    P = slice((0 * 6 + 1) * n2, (0 * 6 + 2) * n2)
    vP = m.vP[P]
    S = st.reshape(6, m.npts)
    Em, Hm = fEN.reshape(3, -1)[:, P], fHN.reshape(3, -1)[:, P]
    Ep, Hp = S[0:3, vP], S[3:6, vP]
    n = np.array([1.0, 0.0, 0.0])[:, None]
    Z1, Z2 = 1.0, math.sqrt(2.0 / 3.0)
    # interface state from the characteristics (E* - Z1 nxH* = E- - Z1 nxH-, E* + Z2 nxH* = E+ + Z2 nxH+)
    nxHm, nxHp = np.cross(n.T, Hm.T).T, np.cross(n.T, Hp.T).T
    tan = lambda v: v - n * np.sum(n * v, axis=0)
    nxHs = (tan(Ep) - tan(Em) + Z2 * nxHp + Z1 * nxHm) / (Z1 + Z2)
    Es_t = (Z2 * tan(Em) + Z1 * tan(Ep) + Z1 * Z2 * (nxHp - nxHm)) / (Z1 + Z2)
    np.testing.assert_allclose(fl[3:6, P], nxHs - nxHm, atol=1e-13)          # F_E = n x (H* - H-)
    nxEs = np.cross(n.T, Es_t.T).T
    nxEm = np.cross(n.T, Em.T).T
    np.testing.assert_allclose(fl[0:3, P], -(nxEs - nxEm), atol=1e-13)       # F_H = -n x (E* - E-)


def test_self_halo_bitwise_equals_periodic():
    m0 = po.Mesh.box(4, (4, 2, 2), UNIT, (1, 0, 0))
    m1 = po.Mesh.box(4, (4, 2, 2), UNIT, (1, 0, 0))
    rng = np.random.default_rng(7)
    eps = np.exp(0.3 * rng.standard_normal(m0.nelt)); mu = np.exp(0.3 * rng.standard_normal(m0.nelt))
    m0.set_materials(eps, mu); m1.set_materials(eps, mu)
    nh = m1.self_halo_x()
    assert nh == 2 * 2 * 2 * 16
    x, y, z = m0.x, m0.y, m0.zc
    st = an.pack([np.sin(2 * math.pi * x), y, z * x, np.cos(2 * math.pi * x) * y, x, z])
    dt = m0.auto_dt()
    a = _run(m0, st, dt, 3)
    b = _run(m1, st, dt, 3)
    assert np.array_equal(a, b)


def test_numpy_lift_equals_oracle_lift():
    """The point-by-point lift the kernel-level tests use as their reference for
    foreign face layouts equals the oracle's or_add_flux on the element-major
    layout (the E rows come from the flux's H-ordered second half)."""
    from gpu_util import lift_numpy

    o = po.Mesh.box(4, (2, 2, 1), UNIT, (1, 0, 0), jitter=0.1, seed=3)
    rng = np.random.default_rng(4)
    fl = rng.standard_normal(6 * o.tfp)
    rhs0 = rng.standard_normal(6 * o.npts)
    want = po.add_flux(o, fl, rhs0)
    got = lift_numpy(o.vM, o.dA, fl, rhs0, o.tfp, o.npts)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14 * np.abs(want).max())
