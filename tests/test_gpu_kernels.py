"""Kernel-level parity: every reference kernel's counterpart, called through the
C ABI on oracle-generated inputs, against the oracle's restatement of that kernel
(dg_kernels.cu:41-453).  Integer/byte movement is bit-exact; FP64 arithmetic
within the stated tolerance (the reference's lift accumulates with FP64
atomics, dg_kernels.cu:307-335, so its own last bits vary from run to run)."""
import numpy as np
import pytest

import nekwave_amd as nw
import pyoracle as po
from nekwave_amd import DeviceBuffer as DB
from nekwave_amd._lib import check, lib
from gpu_util import lift_numpy

pytestmark = pytest.mark.gpu


def _mesh(media=False, halo=False):
    o = po.Mesh.box(5, (3, 2, 2), (0, 1.5, 0, 1, 0, 1), (1, 0, 0), jitter=0.1, seed=4)
    if media:
        rng = np.random.default_rng(2)
        o.set_materials(rng.uniform(1, 4, o.nelt), rng.uniform(0.5, 2, o.nelt))
    if halo:
        o.self_halo_x()
    return o


def _state(o, stride):
    rng = np.random.default_rng(11)
    return rng.standard_normal(6 * stride)


def test_volume_curl_parity():
    o = _mesh()
    rng = np.random.default_rng(0)
    u = [rng.standard_normal(o.npts) for _ in range(3)]
    ref = po.volume_curl(o, *u, -1.0)
    bufs = {k: DB.of(getattr(o, k)) for k in ("w3", "jac", "rx", "sx", "tx", "ry", "sy", "ty", "rz", "sz", "tz", "D")}
    du = [DB.of(a) for a in u]
    dw = [DB(8 * o.npts) for _ in range(3)]
    b = bufs
    check(lib().nw_launch_volume_curl(du[0].ptr, du[1].ptr, du[2].ptr, dw[0].ptr, dw[1].ptr, dw[2].ptr,
                                      b["D"].ptr, b["w3"].ptr, b["jac"].ptr, b["rx"].ptr, b["sx"].ptr, b["tx"].ptr,
                                      b["ry"].ptr, b["sy"].ptr, b["ty"].ptr, b["rz"].ptr, b["sz"].ptr, b["tz"].ptr,
                                      -1.0, o.n1, o.nelt, None))
    check(lib().nw_dev_sync())
    for c in range(3):
        got = dw[c].download(np.float64, o.npts)
        scale = np.abs(ref[c]).max()
        np.testing.assert_allclose(got, ref[c], rtol=0, atol=1e-13 * scale)


@pytest.mark.parametrize("media", [False, True])
@pytest.mark.parametrize("halo", [False, True])
def test_restrict_flux_lift_parity(media, halo):
    o = _mesh(media, halo)
    S = o.npts + o.nhalo
    st = _state(o, S)
    T = o.tfp
    fEN, fHN = po.restrict(o, st, S)
    fl = po.flux(o, st, fEN, fHN, 1.0, S)
    rhs0 = np.random.default_rng(5).standard_normal(6 * o.npts)
    rhs = po.add_flux(o, fl, rhs0)
    d_st = DB.of(st)
    d_vM, d_vP, d_bc = DB.of(o.vM), DB.of(o.vP), DB.of(o.bc)
    d_n = [DB.of(o.nx), DB.of(o.ny), DB.of(o.nz)]
    d_dA = DB.of(o.dA)
    d_fE, d_fH, d_fl = DB(8 * 3 * T), DB(8 * 3 * T), DB(8 * 6 * T)
    check(lib().nw_launch_restrict_faces(d_st.ptr, d_fE.ptr, d_fH.ptr, d_vM.ptr, T, S, None))
    check(lib().nw_dev_sync())
    assert np.array_equal(d_fE.download(np.float64, 3 * T), fEN)
    assert np.array_equal(d_fH.download(np.float64, 3 * T), fHN)
    eps = DB.of(o.eps) if media else None
    mu = DB.of(o.mu) if media else None
    ge = DB.of(o.ghost_elem) if (media and halo) else None
    check(lib().nw_launch_compute_flux(d_st.ptr, d_fE.ptr, d_fH.ptr, d_vP.ptr, d_bc.ptr, d_n[0].ptr, d_n[1].ptr,
                                       d_n[2].ptr, d_fl.ptr, 1.0, T, S, d_vM.ptr,
                                       eps.ptr if eps else None, mu.ptr if mu else None,
                                       ge.ptr if ge else None, o.np, o.npts, None))
    check(lib().nw_dev_sync())
    got = d_fl.download(np.float64, 6 * T)
    np.testing.assert_allclose(got, fl, rtol=0, atol=1e-14 * np.abs(fl).max())
    d_rhs = DB.of(rhs0)
    d_fl.upload(fl)  # lift the oracle flux so the lift is checked on its own
    # the reference's argument list (dg_kernels.hpp:46-50)
    check(lib().nw_launch_add_flux(d_fl.ptr, d_vM.ptr, d_dA.ptr, d_rhs.ptr, T, o.npts, None))
    check(lib().nw_dev_sync())
    np.testing.assert_allclose(d_rhs.download(np.float64, 6 * o.npts), rhs, rtol=0, atol=1e-14 * np.abs(rhs).max())


def test_add_flux_with_a_foreign_face_layout():
    """volIdxMinus that departs from the element-major layout (permuted points,
    points sent to other nodes) and sizes that admit no n1: the points the layout
    does not explain are added as the reference adds them."""
    o = _mesh(False, False)
    T, P = o.tfp, o.npts
    rng = np.random.default_rng(21)
    fl = rng.standard_normal(6 * T)
    dA = rng.uniform(0.5, 1.5, T)
    rhs0 = rng.standard_normal(6 * P)
    vM = o.vM.copy()
    sel = rng.choice(T, T // 5, replace=False)
    vM[sel] = rng.integers(0, P, sel.size).astype(vM.dtype)
    want = lift_numpy(vM, dA, fl, rhs0, T, P)
    d_fl, d_dA, d_vM, d_rhs = DB.of(fl), DB.of(dA), DB.of(vM), DB.of(rhs0)
    check(lib().nw_launch_add_flux(d_fl.ptr, d_vM.ptr, d_dA.ptr, d_rhs.ptr, T, P, None))
    check(lib().nw_dev_sync())
    np.testing.assert_allclose(d_rhs.download(np.float64, 6 * P), want, rtol=0, atol=1e-13 * np.abs(want).max())
    # a subset of the face points (no n1 fits the sizes): every point scattered
    T2 = T - 7
    want2 = lift_numpy(vM[:T2], dA[:T2], np.concatenate([fl[c * T:c * T + T2] for c in range(6)]), rhs0, T2, P)
    d_fl2 = DB.of(np.concatenate([fl[c * T:c * T + T2] for c in range(6)]))
    d_rhs2 = DB.of(rhs0)
    check(lib().nw_launch_add_flux(d_fl2.ptr, d_vM.ptr, d_dA.ptr, d_rhs2.ptr, T2, P, None))
    check(lib().nw_dev_sync())
    np.testing.assert_allclose(d_rhs2.download(np.float64, 6 * P), want2, rtol=0, atol=1e-13 * np.abs(want2).max())


@pytest.mark.parametrize("media", [False, True])
def test_inv_mass_and_lsrk_parity(media):
    o = _mesh(media)
    rng = np.random.default_rng(9)
    rhs = rng.standard_normal(6 * o.npts)
    ref = po.inv_mass(o, rhs)
    d_rhs, d_jac, d_w3 = DB.of(rhs), DB.of(o.jac), DB.of(o.w3)
    eps = DB.of(o.eps) if media else None
    mu = DB.of(o.mu) if media else None
    check(lib().nw_launch_inv_mass(d_rhs.ptr, d_jac.ptr, d_w3.ptr, o.nelt, o.np, o.npts,
                                   eps.ptr if eps else None, mu.ptr if mu else None, None))
    check(lib().nw_dev_sync())
    np.testing.assert_allclose(d_rhs.download(np.float64, 6 * o.npts), ref, rtol=1e-15, atol=0)
    a, b = po.rk_coeffs()
    st = rng.standard_normal(6 * o.npts)
    k = rng.standard_normal(6 * o.npts)
    st_ref, k_ref = po.lsrk_update(o, st, k, ref, a[2], b[2], 0.01)
    d_st, d_k = DB.of(st), DB.of(k)
    check(lib().nw_launch_lsrk45_update(d_st.ptr, d_k.ptr, d_rhs.ptr, a[2], b[2], 0.01, o.npts, o.npts, None))
    check(lib().nw_dev_sync())
    # the device contracts a*k + dt*rhs into one FMA: compare to the size of the terms
    kterm = np.abs(a[2] * k) + np.abs(0.01 * ref)
    gk = d_k.download(np.float64, 6 * o.npts)
    assert np.all(np.abs(gk - k_ref) <= 1e-15 * kterm)
    sterm = np.abs(st) + np.abs(b[2]) * (np.abs(k_ref) + kterm)  # k's own rounding feeds u
    assert np.all(np.abs(d_st.download(np.float64, 6 * o.npts) - st_ref) <= 1e-15 * sterm)


def test_pack_unpack_bit_exact():
    o = _mesh(halo=True)
    S = o.npts + o.nhalo
    st = _state(o, S)
    M = o.nhalo
    buf = np.zeros(6 * M)
    po.lib().or_pack_halo(o._p, po.dptr(st), S, po.dptr(buf))
    st2 = st.copy()
    po.lib().or_unpack_halo(o._p, po.dptr(buf), po.dptr(st2), S)
    off = np.arange(M, dtype=np.int32)
    sizes = np.full(M, M, dtype=np.int32)
    d_st, d_buf = DB.of(st), DB(8 * 6 * M)
    d_vol, d_off, d_sz = DB.of(o.send_vol), DB.of(off), DB.of(sizes)
    check(lib().nw_launch_pack_halo(d_st.ptr, d_buf.ptr, d_vol.ptr, d_off.ptr, d_sz.ptr, M, S, None))
    check(lib().nw_dev_sync())
    assert np.array_equal(d_buf.download(np.float64, 6 * M), buf)
    check(lib().nw_launch_unpack_halo(d_buf.ptr, d_st.ptr, d_off.ptr, d_sz.ptr, M, o.npts, S, None))
    check(lib().nw_dev_sync())
    assert np.array_equal(d_st.download(np.float64, 6 * S), st2)


def test_device_is_gfx950():
    info = nw.device_info(0)
    assert info["count"] >= 1
    assert info["arch"].startswith("gfx950"), info


def _diag_call(st, S, o, media, out):
    d_st, d_jac, d_w = DB.of(st), DB.of(o.jac), DB.of(o.w)
    eps = DB.of(o.eps) if media else None
    mu = DB.of(o.mu) if media else None
    check(lib().nw_launch_diagnostics(d_st.ptr, S, d_jac.ptr, d_w.ptr, eps.ptr if eps else None,
                                      mu.ptr if mu else None, o.n1, o.nelt, out.ptr, None))
    check(lib().nw_dev_sync())
    return out.download(np.float64, 3)


@pytest.mark.parametrize("media", [False, True])
@pytest.mark.parametrize("halo", [False, True])
def test_diagnostics_launch_matches_oracle(media, halo):
    """nw_launch_diagnostics (Case::computeTotalEnergy / getMaxE / getMaxH,
    case.cpp:733-794, on a device state with ghost stride) against the oracle's
    energy (1e-13 relative: only the summation order differs) and a numpy brute
    force of max |E|, max |H| over the npts nodes (1e-15 relative)."""
    o = _mesh(media, halo)
    S = o.npts + o.nhalo
    st = _state(o, S)
    n = o.npts
    got = _diag_call(st, S, o, media, DB(24))
    want_e = o.energy(st, S)
    mag = [np.sqrt(sum(st[c * S:c * S + n] ** 2 for c in cs)).max() for cs in ((0, 1, 2), (3, 4, 5))]
    assert abs(got[0] - want_e) <= 1e-13 * want_e, (got[0], want_e)
    assert abs(got[1] - mag[0]) <= 1e-15 * mag[0] and abs(got[2] - mag[1]) <= 1e-15 * mag[1], (got, mag)
    # distinct E and H: a component-offset slip would swap or mix them
    assert abs(mag[0] - mag[1]) > 1e-6 * mag[0]


def test_diagnostics_launch_tm110_closed_form():
    """TM110 on the PEC unit cube at t = 0: max |E| = 1 at the nodes x = y = 0.5
    (element faces of the 4^3 mesh, where GLL end nodes sit), max |H| = 0, and
    the energy is the continuous 1/8 up to the quadrature error."""
    o = po.Mesh.box(7, (4, 4, 4), (0, 1, 0, 1, 0, 1))
    x, y = o.x, o.y
    z = np.zeros_like(x)
    st = np.concatenate([z, z, np.sin(np.pi * x) * np.sin(np.pi * y), z, z, z])
    got = _diag_call(st, o.npts, o, False, DB(24))
    assert abs(got[1] - 1.0) <= 1e-15 and got[2] == 0.0
    assert abs(got[0] - 0.125) <= 1e-6, got[0]
