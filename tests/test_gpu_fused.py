"""Fused visibility passes around the gridder (hipng_vis_prologue / hipng_vis_epilogue)
against the oracle's literal pieces: flag mask, the phasor of
visibility/base.py:1248-1264, the polarisation matrices of polarisation.py:240-324; and
invert_ng / predict_ng with an image centre away from the visibility phase centre."""

import ctypes
import math

import numpy
import pytest

from oracle import explicit as ox
from rascil_amd import Direction, PolarisationFrame, create_image, invert_ng, predict_ng, simulation
from rascil_amd.data_models import direction_to_lmn

pytestmark = pytest.mark.gpu


def _obs(pol="linear", nchan=3, ntimes=6, seed=0):
    rng = numpy.random.default_rng(seed)
    ants = simulation.gaussian_disc_stations(20, 300.0, 1000.0, rng)
    ha = numpy.linspace(-1, 1, ntimes) * math.pi / 12
    bvis = simulation.create_blockvisibility(ants, ha, numpy.linspace(1e8, 1.1e8, nchan),
                                             Direction(0.0, math.radians(-45)), PolarisationFrame(pol))
    bvis.vis = rng.normal(size=bvis.vis.shape) + 1j * rng.normal(size=bvis.vis.shape)
    bvis.flags = (rng.uniform(size=bvis.vis.shape) < 0.2).astype(int)
    bvis.imaging_weight = rng.uniform(0.5, 2.0, bvis.vis.shape)
    return bvis


def _uvw_lambda(bvis):
    return bvis.uvw[..., None, :] * (bvis.frequency / 299792458.0)[None, None, :, None]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 2e-7)])
@pytest.mark.parametrize("frames", [("linear", "stokesIQUV"), ("circular", "stokesI"), ("stokesI", "stokesI"),
                                    ("linearnp", "stokesIQ"), ("circularnp", "stokesIV"), ("linearnp", "stokesI"),
                                    ("circularnp", "stokesI")])
def test_prologue_matches_oracle(precision, tol, frames):
    import torch

    from rascil_amd import _lib
    from rascil_amd.ng import _conv

    ipf, opf = frames
    bvis = _obs(pol=ipf)
    nt, nb, nc, nvp = bvis.vis.shape
    nop = PolarisationFrame(opf).npol
    l, m, nm1 = direction_to_lmn(Direction(0.01, math.radians(-44.5)), bvis.phasecentre)
    ph = ox.phasor(_uvw_lambda(bvis), l, m, nm1)
    x = bvis.vis * (1 - bvis.flags) * numpy.conj(ph)[..., None]
    ref = ox.convert_pol(x, ipf, opf, polaxis=-1) if ipf != opf else x
    conv, _ = _conv(PolarisationFrame(ipf), PolarisationFrame(opf))
    dev = torch.device("cuda", 0)
    lib = _lib.load()
    n = nt * nb
    v = torch.tensor(bvis.vis, device=dev)
    f = torch.tensor(bvis.flags.astype(numpy.int32), device=dev)
    w = torch.tensor(bvis.imaging_weight, device=dev)
    u = torch.tensor(bvis.uvw.reshape(n, 3), device=dev)
    cdt, rdt = (torch.complex128, torch.float64) if precision == "fp64" else (torch.complex64, torch.float32)
    ms = torch.empty((n, nc, nop), dtype=cdt, device=dev)
    wg = torch.empty((n, nc, nop), dtype=rdt, device=dev)
    nz = torch.full((nc, nop), 7, dtype=torch.int32, device=dev)
    fr = numpy.ascontiguousarray(bvis.frequency)
    lmn = numpy.array([l, m, nm1])
    pd = ctypes.POINTER(ctypes.c_double)
    _lib.check(lib.hipng_vis_prologue(
        ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(w.data_ptr()),
        ctypes.c_void_p(u.data_ptr()), fr.ctypes.data_as(pd), n, nc, nvp, nop, conv.ctypes.data_as(pd),
        lmn.ctypes.data_as(pd), _lib.PRECISIONS[precision], ctypes.c_void_p(ms.data_ptr()),
        ctypes.c_void_p(wg.data_ptr()), ctypes.c_void_p(nz.data_ptr()), None))
    torch.cuda.synchronize()
    got = ms.cpu().numpy().reshape(ref.shape)
    assert numpy.max(numpy.abs(got - ref)) <= tol * numpy.max(numpy.abs(ref))
    refw = (bvis.imaging_weight * (1 - bvis.flags))[..., :nop]
    numpy.testing.assert_allclose(wg.cpu().numpy().reshape(refw.shape), refw, rtol=tol)
    numpy.testing.assert_array_equal(nz.cpu().numpy(), (numpy.abs(ref).reshape(-1, nc, nop).max(axis=0) > 0))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-15)])
@pytest.mark.parametrize("frames", [("stokesIQUV", "linear"), ("stokesIQUV", "circular"), ("stokesI", "stokesI"),
                                    ("stokesIQ", "linearnp"), ("stokesIV", "circularnp")])
def test_epilogue_matches_oracle(precision, tol, frames):
    import torch

    from rascil_amd import _lib
    from rascil_amd.ng import _conv

    ipf, opf = frames
    bvis = _obs(pol=opf)
    nt, nb, nc, nvp = bvis.vis.shape
    nip = PolarisationFrame(ipf).npol
    rng = numpy.random.default_rng(5)
    vist = (rng.normal(size=(nt, nb, nc, nip)) + 1j * rng.normal(size=(nt, nb, nc, nip)))
    vist = vist.astype(numpy.complex128 if precision == "fp64" else numpy.complex64)
    l, m, nm1 = direction_to_lmn(Direction(-0.02, math.radians(-45.3)), bvis.phasecentre)
    ph = ox.phasor(_uvw_lambda(bvis), l, m, nm1)
    x = vist.astype(numpy.complex128)
    ref = (ox.convert_pol(x, ipf, opf, polaxis=-1) if ipf != opf else x) * ph[..., None]
    conv, _ = _conv(PolarisationFrame(ipf), PolarisationFrame(opf))
    dev = torch.device("cuda", 0)
    lib = _lib.load()
    n = nt * nb
    vt = torch.tensor(vist, device=dev)
    u = torch.tensor(bvis.uvw.reshape(n, 3), device=dev)
    out = torch.empty((n, nc, nvp), dtype=torch.complex128, device=dev)
    fr = numpy.ascontiguousarray(bvis.frequency)
    lmn = numpy.array([l, m, nm1])
    pd = ctypes.POINTER(ctypes.c_double)
    _lib.check(lib.hipng_vis_epilogue(
        ctypes.c_void_p(vt.data_ptr()), ctypes.c_void_p(u.data_ptr()), fr.ctypes.data_as(pd), n, nc, nip, nvp,
        conv.ctypes.data_as(pd), lmn.ctypes.data_as(pd), _lib.PRECISIONS[precision],
        ctypes.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(ref.shape)
    assert numpy.max(numpy.abs(got - ref)) <= 1e-13 * numpy.max(numpy.abs(ref))


def test_predict_and_invert_with_shifted_image_centre():
    """A unit pixel at the centre of an image whose centre is offset from the visibility
    phase centre predicts the DFT of a unit source at that centre (phasor of
    visibility/base.py:1248-1264), and inverting those visibilities with the same image puts
    the source back at its centre with unit flux."""
    bvis = _obs(pol="stokesI", nchan=2, ntimes=10)
    bvis.flags[...] = 0
    bvis.imaging_weight[...] = 1.0
    centre = Direction(0.004, math.radians(-44.8))
    N, cell = 128, 4e-4
    model = create_image(N, cell, centre, [1.05e8], channel_bandwidth=[2e7])
    model.pixels[0, 0, N // 2, N // 2] = 1.0
    pred = predict_ng(bvis, model, epsilon=1e-12)
    l, m, nm1 = direction_to_lmn(centre, bvis.phasecentre)
    ref = ox.phasor(_uvw_lambda(bvis), l, m, nm1)
    assert numpy.max(numpy.abs(pred.vis[..., 0] - ref)) < 1e-9
    assert pred.phasecentre is centre
    bvis2 = _obs(pol="stokesI", nchan=2, ntimes=10)
    bvis2.flags[...] = 0
    bvis2.imaging_weight[...] = 1.0
    bvis2.weight[...] = 1.0
    bvis2.vis[..., 0] = ref
    im, _ = invert_ng(bvis2, model, epsilon=1e-12)
    p = im.pixels[0, 0]
    assert numpy.unravel_index(numpy.argmax(p), p.shape) == (N // 2, N // 2)
    assert p[N // 2, N // 2] == pytest.approx(1.0, abs=1e-9)


def test_invert_with_flags_equals_adjoint_of_masked_data():
    """Flagged visibilities drop out of both the image and sumwt (memory_data_models.py:1502-1524)."""
    bvis = _obs(pol="stokesI", nchan=2, ntimes=4)
    N, cell = 64, 8e-4
    model = create_image(N, cell, bvis.phasecentre, [1.05e8])
    im, sumwt = invert_ng(bvis, model, normalise=False, epsilon=1e-12)
    keep = 1 - bvis.flags
    uvw = bvis.uvw.reshape(-1, 3) * [-1, 1, -1]
    ms = (bvis.vis * keep).reshape(-1, 2, 1)[..., 0]
    wgt = (bvis.imaging_weight * keep).reshape(-1, 2, 1)[..., 0]
    ref = ox.explicit_ms2dirty(uvw, bvis.frequency, ms, wgt, N, N, cell, cell, True).T
    assert numpy.max(numpy.abs(im.pixels[0, 0] - ref)) < 1e-10 * numpy.max(numpy.abs(ref))
    assert sumwt[0, 0] == pytest.approx(float(numpy.sum(bvis.imaging_weight * keep)), rel=1e-14)
