"""nw_launch_diagnostics_noalloc: the diagnostics reduction with its partial
sums in the library's fixed per-device array instead of an allocation.  Its
results must equal nw_launch_diagnostics' bit for bit (same kernels, same grid)
on two calls in a row, the second with a smaller state (the array keeps the
first call's partials; every slot the second grid reads it writes first), and
match the oracle's energy and a brute-force max |E|, max |H|."""
import numpy as np
import pytest

import pyoracle as po
from nekwave_amd import DeviceBuffer as DB
from nekwave_amd._lib import NW_ERR_ARG, check, lib

pytestmark = pytest.mark.gpu


def _case(nel, n1, seed):
    o = po.Mesh.box(n1, nel, (0, 1, 0, 1.5, 0, 1), (0, 0, 0), jitter=0.1, seed=seed)
    rng = np.random.default_rng(seed)
    o.set_materials(rng.uniform(1, 4, o.nelt), rng.uniform(0.5, 2, o.nelt))
    S = o.npts + 5  # a stride past npts, as with ghost slots
    st = rng.standard_normal(6 * S)
    return o, S, st


def _call(fn, o, S, st):
    d = [DB.of(a) for a in (st, o.jac, o.w, o.eps, o.mu)]
    out = DB(24)
    check(fn(d[0].ptr, S, d[1].ptr, d[2].ptr, d[3].ptr, d[4].ptr, o.n1, o.nelt, out.ptr, None))
    check(lib().nw_dev_sync())
    return out.download(np.float64, 3)


def test_noalloc_equals_allocating_entry_point_twice():
    # the larger state first (its grid fills all 1024 partial slots), then a smaller one
    for nel, n1, seed in (((14, 10, 9), 6, 3), ((2, 3, 2), 4, 4)):
        o, S, st = _case(nel, n1, seed)
        got = _call(lib().nw_launch_diagnostics_noalloc, o, S, st)
        want = _call(lib().nw_launch_diagnostics, o, S, st)
        assert np.array_equal(got, want), (nel, got, want)
        n = o.npts
        e = o.energy(st, S)
        mag = [np.sqrt(sum(st[c * S:c * S + n] ** 2 for c in cs)).max() for cs in ((0, 1, 2), (3, 4, 5))]
        assert abs(got[0] - e) <= 1e-13 * e, (got[0], e)
        assert abs(got[1] - mag[0]) <= 1e-15 * mag[0] and abs(got[2] - mag[1]) <= 1e-15 * mag[1]


def test_noalloc_rejects_bad_arguments():
    o, S, st = _case((2, 2, 2), 4, 5)
    d = DB.of(st)
    out = DB(24)
    # stride shorter than the state, and a null state: refused before any launch
    assert lib().nw_launch_diagnostics_noalloc(d.ptr, o.npts - 1, d.ptr, d.ptr, None, None, o.n1, o.nelt,
                                               out.ptr, None) == NW_ERR_ARG
    assert lib().nw_launch_diagnostics_noalloc(None, S, d.ptr, d.ptr, None, None, o.n1, o.nelt,
                                               out.ptr, None) == NW_ERR_ARG
