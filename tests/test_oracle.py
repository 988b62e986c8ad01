"""Pins the oracle (oracle/explicit.py) against the reference's known answers and
closed forms before anything is compared with it."""

import numpy
import pytest

from oracle import explicit as ox

C = ox.C_M_S


def _kat_uvw_lambda(ntimes=2, nbl=30, nchan=10, seed=0):
    rng = numpy.random.default_rng(seed)
    return rng.normal(0, 300.0, (ntimes, nbl, nchan, 3))


@pytest.mark.parametrize("ncomp,nchan,maxabs,minabs,rms", [
    # tests/processing_components/test_dft_skycomponent_visibility_kernels.py:83-85
    (100, 10, 12000.0, 1004.987562112086, 4714.611562943335),
    # same file :124-126
    (2, 2, 240.0, 20.099751242241776, 94.29223125886809),
])
def test_dft_known_answers(ncomp, nchan, maxabs, minabs, rms):
    # all components co-located: |V| does not depend on the array configuration
    flux_iquv = numpy.array(nchan * [100.0, 20.0, -10.0, 1.0]).reshape(nchan, 4)
    flux_lin = ox.convert_pol(flux_iquv.astype(complex), "stokesIQUV", "linear", polaxis=1)
    vfl = numpy.array(ncomp * [flux_lin])
    l, m = 0.0143, -0.002
    dc = numpy.array(ncomp * [[l, m, numpy.sqrt(1 - l * l - m * m) - 1]])
    vis = ox.dft_cpu_looped(dc, _kat_uvw_lambda(nchan=nchan), vfl)
    a = numpy.abs(vis)
    numpy.testing.assert_almost_equal(a.max(), maxabs)
    numpy.testing.assert_almost_equal(a.min(), minabs)
    numpy.testing.assert_almost_equal(a.std(), rms)  # qa_visibility "rms" is numpy.std(|vis|)


def test_dft_known_answer_catches_wrong_pol_matrix():
    """A transposed linear matrix (plausible slip) changes the known answer."""
    flux = numpy.array([100.0, 20.0, -10.0, 1.0])
    good = ox.POL_MATRICES[("stokesIQUV", "linear")] @ flux
    bad = ox.POL_MATRICES[("stokesIQUV", "linear")].T @ flux
    assert numpy.abs(good).max() == pytest.approx(120.0)
    assert not numpy.allclose(numpy.abs(good), numpy.abs(bad))


def test_pol_roundtrip_and_inverse():
    for a, b in [("stokesIQUV", "linear"), ("stokesIQUV", "circular")]:
        m = ox.POL_MATRICES[(a, b)] @ ox.POL_MATRICES[(b, a)]
        numpy.testing.assert_allclose(m, numpy.eye(4), atol=1e-15)


def test_unit_pixel_predict_is_one(small_obs):
    """test_imaging.py:274-282: unit pixel at (N/2, N/2) predicts vis == 1 (2d)."""
    uvw, _ = small_obs
    d = numpy.zeros((32, 32))
    d[16, 16] = 1.0
    ms = ox.explicit_dirty2ms(uvw[:200], [1e8], d, None, 1e-3, 1e-3, False)
    numpy.testing.assert_allclose(ms, 1.0, atol=1e-12)
    # with w-stacking the centre pixel has n = 1 and no w phase: still exactly 1
    ms = ox.explicit_dirty2ms(uvw[:200], [1e8], d, None, 1e-3, 1e-3, True)
    numpy.testing.assert_allclose(ms, 1.0, atol=1e-12)


def test_single_visibility_closed_form():
    """dirty(x, y) for one visibility equals wgt Re[V exp(2 pi i (u x + v y - w (n-1)))] / n."""
    uvw = numpy.array([[120.0, -45.0, 33.0]])
    f = 1.5e8
    V = 0.3 - 0.7j
    nx, ny, px, py = 16, 12, 3e-3, 2e-3
    d = ox.explicit_ms2dirty(uvw, [f], numpy.array([[V]]), numpy.array([[2.0]]), nx, ny, px, py, True)
    ix, iy = 3, 10
    x, y = (ix - nx // 2) * px, (iy - ny // 2) * py
    n = numpy.sqrt(1 - x * x - y * y)
    u, v, w = uvw[0] * f / C
    expect = 2.0 * (V * numpy.exp(2j * numpy.pi * (u * x + v * y - w * (n - 1)))).real / n
    assert d[ix, iy] == pytest.approx(expect, rel=1e-13)


def test_adjoint_identity(small_obs):
    """<dirty2ms(I), V>_R == <I, ms2dirty(V)> for the explicit sums."""
    uvw, _ = small_obs
    rng = numpy.random.default_rng(1)
    freq = [1e8, 1.2e8]
    V = rng.normal(size=(300, 2)) + 1j * rng.normal(size=(300, 2))
    I = rng.normal(size=(24, 20))
    for wst in (False, True):
        lhs = numpy.vdot(ox.explicit_dirty2ms(uvw[:300], freq, I, None, 2e-3, 3e-3, wst), V).real
        rhs = numpy.sum(I * ox.explicit_ms2dirty(uvw[:300], freq, V, None, 24, 20, 2e-3, 3e-3, wst))
        assert lhs == pytest.approx(rhs, rel=1e-11)


def test_hermitian_symmetry(small_obs):
    """uvw -> -uvw with V -> conj(V) leaves the dirty image unchanged."""
    uvw, _ = small_obs
    rng = numpy.random.default_rng(2)
    V = rng.normal(size=(200, 1)) + 1j * rng.normal(size=(200, 1))
    a = ox.explicit_ms2dirty(uvw[:200], [1e8], V, None, 20, 20, 2e-3, 2e-3, True)
    b = ox.explicit_ms2dirty(-uvw[:200], [1e8], V.conj(), None, 20, 20, 2e-3, 2e-3, True)
    numpy.testing.assert_allclose(a, b, atol=1e-10 * numpy.abs(a).max())


def test_psf_peak_is_sumwt(small_obs):
    """test_imaging.py:601-626: PSF (ms = 1) peaks at the centre with value sum(wgt)."""
    uvw, _ = small_obs
    rng = numpy.random.default_rng(3)
    wgt = rng.uniform(0.5, 2.0, (400, 1))
    d = ox.explicit_ms2dirty(uvw[:400], [1e8], numpy.ones((400, 1)), wgt, 32, 32, 1e-3, 1e-3, True)
    assert numpy.unravel_index(numpy.argmax(d), d.shape) == (16, 16)
    assert d[16, 16] / wgt.sum() == pytest.approx(1.0, abs=1e-12)


def test_point_source_predict_matches_dft_with_1_over_n(small_obs):
    """A pixel-centred source of flux F predicts the DFT of F/n (ducc0 w-stacking convention)."""
    uvw, _ = small_obs
    N, cell = 32, 2e-3
    ix, iy = 5, 27
    d = numpy.zeros((N, N))
    d[ix, iy] = 2.5
    ms = ox.explicit_dirty2ms(uvw[:100], [1e8], d, None, cell, cell, True)
    x, y = (ix - N // 2) * cell, (iy - N // 2) * cell
    n = numpy.sqrt(1 - x * x - y * y)
    # RASCIL flips u and w before the call: x = -l, y = m
    l, m = -x, y
    ruvw = uvw[:100] * numpy.array([-1, 1, -1])
    uvl = (ruvw * 1e8 / C)[None, :, None, :]
    vis = ox.dft_cpu_looped(numpy.array([[l, m, n - 1]]), uvl, numpy.array([[[2.5 / n]]]))
    numpy.testing.assert_allclose(ms[:, 0], vis[0, :, 0, 0], rtol=1e-11, atol=1e-11)


def test_two_pol_matrices_known_answers():
    """Hand-computed: I = 3, Q = 1 -> XX = I + Q = 4, YY = I - Q = 2 (TMS 4.58 restricted to
    the parallel hands); I = 3, V = -1 -> RR = I + V = 2, LL = I - V = 4 (TMS 4.59); and
    back, and to Stokes I = (XX + YY) / 2."""
    iq = numpy.array([[3.0, 1.0]])
    xy = ox.convert_pol(iq, "stokesIQ", "linearnp")
    numpy.testing.assert_array_equal(xy, [[4.0, 2.0]])
    numpy.testing.assert_array_equal(ox.convert_pol(xy, "linearnp", "stokesIQ"), iq)
    numpy.testing.assert_array_equal(ox.convert_pol(xy, "linearnp", "stokesI"), [[3.0]])
    iv = numpy.array([[3.0, -1.0]])
    rl = ox.convert_pol(iv, "stokesIV", "circularnp")
    numpy.testing.assert_array_equal(rl, [[2.0, 4.0]])
    numpy.testing.assert_array_equal(ox.convert_pol(rl, "circularnp", "stokesIV"), iv)
    numpy.testing.assert_array_equal(ox.convert_pol(rl, "circularnp", "stokesI"), [[3.0]])


def test_two_pol_matrices_are_the_parallel_hands_of_the_four_pol_ones():
    """The 2x2 conversions equal the 4x4 ones restricted to the parallel-hand correlations
    (XX, YY: rows / columns 0, 3) and the Stokes pair the frame carries (I, Q: 0, 1 for
    linear; I, V: 0, 3 for circular). The 4x4 linear matrix is itself pinned by the DFT
    known answers above, so a transposed or sign-flipped 2x2 fails here."""
    M = ox.POL_MATRICES
    numpy.testing.assert_array_equal(M[("stokesIQ", "linearnp")],
                                     M[("stokesIQUV", "linear")][numpy.ix_([0, 3], [0, 1])])
    numpy.testing.assert_array_equal(M[("linearnp", "stokesIQ")],
                                     M[("linear", "stokesIQUV")][numpy.ix_([0, 1], [0, 3])])
    numpy.testing.assert_array_equal(M[("stokesIV", "circularnp")],
                                     M[("stokesIQUV", "circular")][numpy.ix_([0, 3], [0, 3])])
    numpy.testing.assert_array_equal(M[("circularnp", "stokesIV")],
                                     M[("circular", "stokesIQUV")][numpy.ix_([0, 3], [0, 3])])
    for a, b in [("stokesIQ", "linearnp"), ("stokesIV", "circularnp")]:
        numpy.testing.assert_allclose(M[(b, a)] @ M[(a, b)], numpy.eye(2), atol=0)



@pytest.mark.parametrize("wst", [False, True])
def test_probe_pixel_sum_equals_the_whole_image_sum(wst):
    # explicit_ms2dirty_at is the whole-image sum read at chosen pixels
    rng = numpy.random.default_rng(11)
    uvw = rng.normal(0.0, 200.0, (7, 3))
    freq = numpy.array([1.0e8, 1.3e8])
    ms = rng.normal(size=(7, 2)) + 1j * rng.normal(size=(7, 2))
    wgt = rng.uniform(0.5, 1.5, (7, 2))
    nx, ny, px, py = 24, 20, 2e-2, 3e-2
    full = ox.explicit_ms2dirty(uvw, freq, ms, wgt, nx, ny, px, py, wst)
    ix, iy = numpy.meshgrid(numpy.arange(nx), numpy.arange(ny), indexing="ij")
    at = ox.explicit_ms2dirty_at(uvw, freq, ms, wgt, ix.ravel(), iy.ravel(), nx, ny, px, py, wst)
    assert numpy.allclose(at.reshape(nx, ny), full, rtol=0, atol=1e-12 * numpy.abs(full).max())


@pytest.mark.parametrize("wst", [False, True])
def test_sparse_image_sum_equals_the_whole_image_sum(wst):
    # explicit_dirty2ms_sparse is the whole-image forward sum of an image that is zero off the
    # given pixels
    rng = numpy.random.default_rng(12)
    uvw = rng.normal(0.0, 200.0, (7, 3))
    freq = numpy.array([1.0e8, 1.3e8])
    wgt = rng.uniform(0.5, 1.5, (7, 2))
    nx, ny, px, py = 24, 20, 2e-2, 3e-2
    ix = numpy.array([0, 23, 12, 5, 0])
    iy = numpy.array([0, 19, 10, 17, 19])
    val = rng.normal(size=ix.size)
    dirty = numpy.zeros((nx, ny))
    dirty[ix, iy] = val
    full = ox.explicit_dirty2ms(uvw, freq, dirty, wgt, px, py, wst)
    sp = ox.explicit_dirty2ms_sparse(uvw, freq, ix, iy, val, nx, ny, px, py, wgt, wst)
    assert numpy.allclose(sp, full, rtol=0, atol=1e-12 * numpy.abs(full).max())
