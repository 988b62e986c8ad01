"""GPU tests of the RASCIL-level surface (invert_ng / predict_ng / dispatch / DFT) with
the invariants of the reference's own tests (test_imaging.py, test_imaging_ng.py,
test_dft_skycomponent_visibility*.py), restated on synthetic geometry."""

import math

import numpy
import pytest

from conftest import rel_l2
from oracle import explicit as ox
from rascil_amd import (Direction, PolarisationFrame, create_image, dft_skycomponent_visibility,
                        invert_blockvisibility, invert_ng, predict_blockvisibility, predict_ng,
                        simulation)
from rascil_amd.dft import dft_kernel

pytestmark = pytest.mark.gpu


def _obs(pol="stokesI", nchan=1, ntimes=8, zerow=False, seed=0):
    rng = numpy.random.default_rng(seed)
    ants = simulation.gaussian_disc_stations(30, 300.0, 1000.0, rng)
    ha = numpy.linspace(-1, 1, ntimes) * math.pi / 12
    freq = numpy.linspace(1e8, 1.1e8, nchan)
    bvis = simulation.create_blockvisibility(ants, ha, freq, Direction(0.0, math.radians(-45)),
                                             PolarisationFrame(pol))
    if zerow:
        bvis.uvw[..., 2] = 0.0
    return bvis


@pytest.mark.parametrize("context", ["2d", "ng"])
def test_predict_unit_point(context):
    """test_imaging.py:274-282: unit pixel at the image centre predicts vis == 1."""
    bvis = _obs(zerow=(context == "2d"))
    model = create_image(128, 5e-4, bvis.phasecentre, [1e8])
    model.pixels[0, 0, 64, 64] = 1.0
    vis = predict_blockvisibility(bvis, model, context=context, epsilon=1e-12)
    assert numpy.max(numpy.abs(vis.vis - 1.0)) < 1e-10


def test_predict_unit_point_iquv():
    """test_imaging.py:284-293: I=1 at the centre -> XX = YY = 1, XY = YX = 0."""
    bvis = _obs(pol="linear", zerow=True)
    model = create_image(128, 5e-4, bvis.phasecentre, [1e8], PolarisationFrame("stokesIQUV"))
    model.pixels[0, 0, 64, 64] = 1.0
    vis = predict_blockvisibility(bvis, model, context="2d", epsilon=1e-12)
    assert numpy.max(numpy.abs(vis.vis[..., 0] - 1.0)) < 1e-10
    assert numpy.max(numpy.abs(vis.vis[..., 1])) < 1e-10
    assert numpy.max(numpy.abs(vis.vis[..., 2])) < 1e-10
    assert numpy.max(numpy.abs(vis.vis[..., 3] - 1.0)) < 1e-10


@pytest.mark.parametrize("image_pol,vis_pol,pix,expect", [
    ("stokesIQ", "linearnp", (1.0, 0.5), (1.5, 0.5)),    # XX = I + Q, YY = I - Q
    ("stokesIV", "circularnp", (1.0, -0.25), (0.75, 1.25)),  # RR = I + V, LL = I - V
])
def test_predict_unit_point_two_pol(image_pol, vis_pol, pix, expect):
    """test_imaging.py:295-318 (IQ, IV): a point at the image centre predicts constant
    parallel-hand visibilities through the 2x2 conversion (polarisation.py:229-231, 379-393)."""
    bvis = _obs(pol=vis_pol, zerow=True)
    model = create_image(128, 5e-4, bvis.phasecentre, [1e8], PolarisationFrame(image_pol))
    model.pixels[0, :, 64, 64] = pix
    vis = predict_blockvisibility(bvis, model, context="2d", epsilon=1e-12)
    assert vis.vis.shape[-1] == 2
    for p in range(2):
        assert numpy.max(numpy.abs(vis.vis[..., p] - expect[p])) < 1e-10


@pytest.mark.parametrize("image_pol,vis_pol,flux", [("stokesIQ", "linearnp", [10.0, 2.0]),
                                                    ("stokesIV", "circularnp", [10.0, -1.5])])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-6), ("mixed", 1e-4)])
def test_invert_recovers_point_two_pol(image_pol, vis_pol, flux, precision, tol):
    """test_imaging_ng.py:290-300 (invert_ng_IQ / _IV): a point source in a 2-pol frame
    appears at its pixel with its flux ratio after the 2x2 conversion into the image frame."""
    from rascil_amd.data_models import Skycomponent, lmn_to_direction

    bvis = _obs(pol=vis_pol, ntimes=12, nchan=2)
    N, cell = 128, 5e-4
    iy, ix = 70, 40
    l, m = -cell * (ix - N // 2), cell * (iy - N // 2)
    sc = Skycomponent(lmn_to_direction(l, m, bvis.phasecentre), bvis.frequency,
                      numpy.array([flux] * 2), PolarisationFrame(image_pol))
    dft_skycomponent_visibility(bvis, sc)
    model = create_image(N, cell, bvis.phasecentre, [1.05e8], PolarisationFrame(image_pol))
    im, sumwt = invert_ng(bvis, model, epsilon=1e-10 if precision == "fp64" else 1e-5, precision=precision)
    p = im.pixels[0]
    assert numpy.unravel_index(numpy.argmax(p[0]), p[0].shape) == (iy, ix)
    n = math.sqrt(1 - l * l - m * m)
    numpy.testing.assert_allclose(p[:, iy, ix] * n, flux, rtol=tol, atol=tol * 10)
    assert sumwt.shape == (1, 2)
    # and the predict of that image returns the parallel-hand visibilities of the source
    bvis2 = _obs(pol=vis_pol, ntimes=12, nchan=2)
    mod = create_image(N, cell, bvis.phasecentre, [1.05e8], PolarisationFrame(image_pol))
    mod.pixels[0, :, iy, ix] = numpy.array(flux) * n  # a pixel P predicts P / n times the phasor
    pred = predict_ng(bvis2, mod, epsilon=1e-10 if precision == "fp64" else 1e-5, precision=precision)
    assert rel_l2(pred.vis, bvis.vis) < (1e-8 if precision == "fp64" else 1e-4)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_psf_peak_is_one(precision, tol):
    """test_imaging.py:601-612: PSF peak == 1 after normalisation."""
    bvis = _obs(nchan=3)
    model = create_image(128, 5e-4, bvis.phasecentre, [1.05e8])
    psf, sumwt = invert_blockvisibility(bvis, model, dopsf=True, epsilon=1e-12 if precision == "fp64" else 1e-5,
                                        precision=precision)
    assert abs(psf.pixels.max() - 1.0) < tol
    assert numpy.unravel_index(numpy.argmax(psf.pixels[0, 0]), (128, 128)) == (64, 64)
    assert sumwt[0, 0] == pytest.approx(bvis.imaging_weight.sum())


def test_predict_matches_dft_of_flux_over_n():
    """Pixel-centred sources with w-stacking: predict_ng == DFT of F/n (ducc0 convention)."""
    bvis = _obs(ntimes=6)
    N, cell = 256, 5e-4
    model = create_image(N, cell, bvis.phasecentre, [1e8])
    pix = [(100, 40, 1.0), (200, 220, 0.5), (128, 128, 0.25), (10, 250, 0.7)]
    comps = []
    for iy, ix, f in pix:
        model.pixels[0, 0, iy, ix] = f
        l = -cell * (ix - N // 2)
        m = cell * (iy - N // 2)
        n = math.sqrt(1 - l * l - m * m)
        comps.append((l, m, n, f))
    vis = predict_ng(bvis, model, epsilon=1e-10)
    dc = numpy.array([[l, m, n - 1] for l, m, n, _ in comps])
    fl = numpy.array([[[f / n]] for l, m, n, f in comps])
    ref = ox.dft_cpu_looped(dc, bvis.uvw_lambda, fl)
    assert rel_l2(vis.vis, ref) < 1e-9


def test_invert_matches_explicit_adjoint():
    """invert_ng (unnormalised) equals the brute-force adjoint with RASCIL's u/w flip and transpose."""
    bvis = _obs(ntimes=4, nchan=2)
    rng = numpy.random.default_rng(4)
    bvis.vis = rng.normal(size=bvis.vis.shape) + 1j * rng.normal(size=bvis.vis.shape)
    N, cell = 64, 8e-4
    model = create_image(N, cell, bvis.phasecentre, [1.05e8])
    im, sumwt = invert_ng(bvis, model, normalise=False, epsilon=1e-12)
    uvw = bvis.uvw.reshape(-1, 3) * [-1, 1, -1]
    ms = bvis.vis.reshape(-1, 2, 1)[..., 0]
    wgt = bvis.imaging_weight.reshape(-1, 2, 1)[..., 0]
    ref = ox.explicit_ms2dirty(uvw, bvis.frequency, ms, wgt, N, N, cell, cell, True).T
    assert rel_l2(im.pixels[0, 0], ref) < 1e-11


def test_invert_recovers_point_position_and_pol_ratio():
    """test_imaging_ng.py:139-167 style: a point source appears at its pixel with its flux."""
    bvis = _obs(pol="linear", ntimes=12, nchan=2)
    N, cell = 128, 5e-4
    iy, ix = 80, 50
    l, m = -cell * (ix - N // 2), cell * (iy - N // 2)
    from rascil_amd.data_models import Skycomponent, lmn_to_direction

    sc = Skycomponent(lmn_to_direction(l, m, bvis.phasecentre), bvis.frequency,
                      numpy.array([[10.0, 1.0, -2.0, 0.5]] * 2), PolarisationFrame("stokesIQUV"))
    dft_skycomponent_visibility(bvis, sc)
    model = create_image(N, cell, bvis.phasecentre, [1.05e8], PolarisationFrame("stokesIQUV"))
    im, _ = invert_ng(bvis, model, epsilon=1e-10)
    p = im.pixels[0]
    assert numpy.unravel_index(numpy.argmax(p[0]), p[0].shape) == (iy, ix)
    n = math.sqrt(1 - l * l - m * m)
    # the dirty image carries the ducc0 1/n factor; the PSF peak is 1
    numpy.testing.assert_allclose(p[:, iy, ix] * n, [10.0, 1.0, -2.0, 0.5], rtol=1e-6, atol=1e-6)


def test_spectral_invert_and_predict_channels():
    bvis = _obs(nchan=3)
    model = create_image(64, 8e-4, bvis.phasecentre, bvis.frequency)
    model.pixels[:, 0, 32, 32] = [1.0, 2.0, 3.0]
    vis = predict_ng(bvis, model, epsilon=1e-10)
    for c in range(3):
        assert numpy.max(numpy.abs(vis.vis[:, :, c, 0] - (c + 1))) < 1e-8
    im, sumwt = invert_ng(vis, model, epsilon=1e-10)
    assert sumwt.shape == (3, 1)
    for c in range(3):
        assert im.pixels[c, 0, 32, 32] == pytest.approx(c + 1, rel=1e-8)


@pytest.mark.parametrize("vpol", ["linear", "circular", "linearnp", "stokesI"])
def test_dft_then_idft_recovers_component_flux(vpol):
    """test_dft_skycomponent_visibility.py:134-147: the inverse DFT of the DFT of a component
    returns its flux (rtol 1e-10), for IQUV components in linear and circular frames; with
    flags the flagged visibilities and their weights drop out of both sums."""
    from rascil_amd import idft_visibility_skycomponent
    from rascil_amd.data_models import Skycomponent, lmn_to_direction

    bvis = _obs(pol=vpol, nchan=3, ntimes=5)
    f = {"linear": [100.0, 20.0, -10.0, 1.0], "circular": [100.0, 20.0, -10.0, 1.0],
         "linearnp": [100.0, 20.0], "stokesI": [100.0]}[vpol]
    cpol = {"linear": "stokesIQUV", "circular": "stokesIQUV", "linearnp": "stokesIQ", "stokesI": "stokesI"}[vpol]
    flux = numpy.array([f, [0.8 * x for x in f], [0.6 * x for x in f]])
    comps = [Skycomponent(lmn_to_direction(0.01, -0.02, bvis.phasecentre), bvis.frequency, flux,
                          PolarisationFrame(cpol)),
             Skycomponent(lmn_to_direction(-0.03, 0.005, bvis.phasecentre), bvis.frequency, 0.5 * flux,
                          PolarisationFrame(cpol))]
    dft_skycomponent_visibility(bvis, comps[0])
    rcomp, weights = idft_visibility_skycomponent(bvis, comps[0])
    numpy.testing.assert_allclose(numpy.real(rcomp[0].flux), flux, rtol=1e-10)
    assert weights[0].shape == (3, len(f))
    numpy.testing.assert_allclose(weights[0], bvis.weight.sum(axis=(0, 1)), rtol=1e-13)
    # two components: each returned flux equals the oracle's weighted mean of vis x conj phasor
    rng = numpy.random.default_rng(2)
    bvis.flags = (rng.uniform(size=bvis.vis.shape) < 0.3).astype(int)
    bvis.weight = rng.uniform(0.5, 2.0, bvis.vis.shape)
    rc, ws = idft_visibility_skycomponent(bvis, comps)
    fw = bvis.weight * (1 - bvis.flags)
    for k, comp in enumerate(comps):
        from rascil_amd.data_models import direction_to_lmn

        l, m, nm1 = direction_to_lmn(comp.direction, bvis.phasecentre)
        ph = numpy.conj(ox.phasor(bvis.uvw_lambda, l, m, nm1))[..., None]
        ref = (fw * bvis.vis * ph).sum(axis=(0, 1)) / fw.sum(axis=(0, 1))
        if cpol != vpol:
            ref = ox.convert_pol(ref, vpol, cpol, polaxis=-1)
        numpy.testing.assert_allclose(rc[k].flux, ref, rtol=1e-11, atol=1e-11)
        numpy.testing.assert_allclose(ws[k], fw.sum(axis=(0, 1)), rtol=1e-13)


def test_dft_interpolates_component_flux_across_channels():
    """dft.py:94-107: a component tabulated on other frequencies is interpolated per
    polarisation with a cubic spline onto the visibility channels."""
    from scipy import interpolate

    from rascil_amd.data_models import Skycomponent, lmn_to_direction

    bvis = _obs(pol="stokesI", nchan=5, ntimes=4)
    cfreq = numpy.linspace(0.95e8, 1.15e8, 6)
    cflux = (2.0 + numpy.sin(cfreq / 1e7))[:, None]
    comp = Skycomponent(lmn_to_direction(0.02, 0.01, bvis.phasecentre), cfreq, cflux, PolarisationFrame("stokesI"))
    dft_skycomponent_visibility(bvis, comp)
    vflux = interpolate.interp1d(cfreq, cflux[:, 0], kind="cubic")(bvis.frequency)
    from rascil_amd.data_models import direction_to_lmn

    l, m, nm1 = direction_to_lmn(comp.direction, bvis.phasecentre)
    ref = ox.dft_cpu_looped(numpy.array([[l, m, nm1]]), bvis.uvw_lambda, vflux[None, :, None].astype(complex))
    numpy.testing.assert_allclose(bvis.vis, ref, rtol=1e-12, atol=1e-12)


def test_gpu_dft_matches_oracle_and_known_answer():
    rng = numpy.random.default_rng(0)
    uvl = rng.normal(0, 300, (2, 50, 10, 3))
    nchan = 10
    flux_iquv = numpy.array(nchan * [100.0, 20.0, -10.0, 1.0]).reshape(nchan, 4)
    fl = ox.convert_pol(flux_iquv.astype(complex), "stokesIQUV", "linear", polaxis=1)
    dc = numpy.array(100 * [[0.0143, -0.002, math.sqrt(1 - 0.0143 ** 2 - 0.002 ** 2) - 1]])
    vis = dft_kernel(dc, numpy.array(100 * [fl]), uvl)
    a = numpy.abs(vis)
    numpy.testing.assert_almost_equal(a.max(), 12000.0)
    numpy.testing.assert_almost_equal(a.min(), 1004.987562112086)
    numpy.testing.assert_almost_equal(a.std(), 4714.611562943335)
    dcr = numpy.column_stack([rng.uniform(-0.05, 0.05, (7, 2)), numpy.zeros(7)])
    dcr[:, 2] = numpy.sqrt(1 - dcr[:, 0] ** 2 - dcr[:, 1] ** 2) - 1
    flr = rng.normal(size=(7, 10, 4)) + 1j * rng.normal(size=(7, 10, 4))
    numpy.testing.assert_allclose(dft_kernel(dcr, flr, uvl), ox.dft_cpu_looped(dcr, uvl, flr),
                                  rtol=0, atol=1e-11)
