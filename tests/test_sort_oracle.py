"""CPU tests of the bucket-sort oracle: keys follow the coordinate contract and the
counting sort is sorted and stable (the properties the GPU sort is then held to)."""

import math
from fractions import Fraction

import numpy
import pytest

from oracle import sortref
from rascil_amd import wgridder


def fma(a, b, c):
    """Correctly rounded a*b+c (exact rational arithmetic, one rounding)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _plan_and_scales(uvw, freq, wst, prec="fp64"):
    c = 299792458.0
    ul = numpy.abs(uvw).max(axis=0) * max(freq) / c
    wmin = numpy.abs(uvw[:, 2]).min() * min(freq) / c
    p = wgridder.plan_params(ul[0], ul[1], wmin, ul[2], 128, 128, 6e-4, 6e-4, 1e-5, wst, prec)
    sx, sy, sw = wgridder.channel_scales(p, freq)
    return p, p.as_dict(), sx, sy, sw


@pytest.mark.parametrize("wst", [False, True])
def test_oracle_sort_sorted_stable_and_keys(small_obs, wst):
    uvw, _ = small_obs
    freq = numpy.array([1.0e8, 1.07e8, 1.31e8])
    p, pd, sx, sy, sw = _plan_and_scales(uvw, freq, wst)
    ks, perm = sortref.bucket_sort(uvw, sx, sy, sw, pd)
    n = uvw.shape[0] * 3
    assert numpy.all(numpy.diff(ks.astype(numpy.int64)) >= 0)
    assert numpy.array_equal(numpy.sort(perm), numpy.arange(n))
    # independent recomputation of a few keys from the plan (pins the key formula)
    for i in (0, 7, n // 2, n - 1):
        r, ch = divmod(int(perm[i]), 3)
        u, v, w = uvw[r]
        if w < 0:
            u, v, w = -u, -v, -w
        ax = (math.floor(fma(u, sx[ch], pd["gx"] / 2) - p.W / 2) + 1) % pd["gx"]
        ay = (math.floor(fma(v, sy[ch], pd["gy"] / 2) - p.W / 2) + 1) % pd["gy"]
        s = 0
        if wst:
            s = math.floor(fma(w, sw[ch], -(p.w0 / p.dw)) - p.W / 2) + 1
            assert 0 <= s <= p.nplanes - p.W
        key = s * p.ntx * p.nty + (ax // p.tile) * p.nty + ay // p.tile
        assert key == ks[i]
    # within a bucket: ordered by in-tile origin digit, stable among equal digits
    rows, chans = numpy.divmod(perm.astype(numpy.int64), 3)
    flip = uvw[rows, 2] < 0
    u = numpy.where(flip, -uvw[rows, 0], uvw[rows, 0])
    v = numpy.where(flip, -uvw[rows, 1], uvw[rows, 1])
    ax = (numpy.floor(u * sx[chans] + pd["gx"] / 2 - p.W / 2).astype(numpy.int64) + 1) % pd["gx"]
    ay = (numpy.floor(v * sy[chans] + pd["gy"] / 2 - p.W / 2).astype(numpy.int64) + 1) % pd["gy"]
    d = (ax % p.tile) * p.tile + ay % p.tile  # unfused multiply-add: exact away from cell edges
    same_key = ks[1:] == ks[:-1]
    assert numpy.all(d[1:][same_key] >= d[:-1][same_key])
    same = same_key & (d[1:] == d[:-1])
    assert numpy.all(perm[1:][same] > perm[:-1][same])  # stable
