"""Imaging weights: oracle pinned by closed forms (CPU), GPU kernel against the oracle and
the reference's PSF invariant under uniform / robust weighting (GPU)."""

import math

import numpy
import pytest

from oracle import weighting as ow
from rascil_amd import Direction, PolarisationFrame, create_image, simulation


def _obs(nchan=2, pol="stokesI", seed=0):
    rng = numpy.random.default_rng(seed)
    ants = simulation.gaussian_disc_stations(25, 300.0, 1000.0, rng)
    ha = numpy.linspace(-1, 1, 6) * math.pi / 12
    return simulation.create_blockvisibility(ants, ha, numpy.linspace(1e8, 1.1e8, nchan),
                                             Direction(0.0, math.radians(-45)), PolarisationFrame(pol))


def _args(bvis, model):
    nt, nb, nc, npol = bvis.vis.shape
    v2i = numpy.round(model.image_acc.wcs.freq_to_pix(bvis.frequency)).astype(int)
    return (bvis.uvw.reshape(-1, 3), bvis.frequency,
            (bvis.imaging_weight * (1 - bvis.flags)).reshape(-1, nc, npol), v2i, model.pixels.shape[0],
            model.pixels.shape[3], model.pixels.shape[2], model.wcs.cdelt[0], model.wcs.cdelt[1])


def test_oracle_uniform_cell_sums_are_constant():
    """Uniform weighting of unit weights: every occupied cell sums to exactly 1."""
    bvis = _obs()
    model = create_image(64, 1e-3, bvis.phasecentre, [1.05e8], channel_bandwidth=[2e7])
    args = _args(bvis, model)
    imwt = numpy.ones_like(args[2])
    imwt[:, :, :] *= (bvis.weight.reshape(imwt.shape) > 0)  # autocorrelations carry weight 0
    new, grid = ow.weight(args[0], args[1], imwt, *args[3:], "uniform")
    pu, pv = ow.cells(args[0], args[1], 64, 64, model.wcs.cdelt[0], model.wcs.cdelt[1])
    cellsum = numpy.zeros((64, 64))
    for ch in range(2):
        numpy.add.at(cellsum, (pv[:, ch], pu[:, ch]), new[:, ch, 0])
    occ = grid[0, 0] > 0
    numpy.testing.assert_allclose(cellsum[occ], 1.0, rtol=1e-12)


def test_oracle_robust_limits():
    bvis = _obs()
    model = create_image(64, 1e-3, bvis.phasecentre, [1.05e8], channel_bandwidth=[2e7])
    args = _args(bvis, model)
    uni, _ = ow.weight(*args, "uniform")
    nat_like, _ = ow.weight(*args, "robust", robustness=12.0)  # f2 -> 0: natural
    uni_like, _ = ow.weight(*args, "robust", robustness=-12.0)  # f2 -> inf: uniform
    numpy.testing.assert_allclose(nat_like, args[2] * (args[2] > 0), rtol=1e-9)
    numpy.testing.assert_allclose(uni_like, uni, rtol=1e-9, atol=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("weighting,rob", [("uniform", 0.0), ("robust", 0.0), ("robust", -1.5), ("robust", 2.0)])
def test_gpu_weighting_matches_oracle(weighting, rob):
    from rascil_amd.weighting import weight_visibility

    bvis = _obs(nchan=3, pol="linear")
    model = create_image(128, 5e-4, bvis.phasecentre, [1.0e8, 1.05e8, 1.1e8], PolarisationFrame("linear"))
    args = _args(bvis, model)
    ref, _ = ow.weight(*args, weighting, robustness=rob)
    out = weight_visibility(bvis, model, weighting=weighting, robustness=rob)
    numpy.testing.assert_allclose(out.imaging_weight.reshape(ref.shape), ref, rtol=1e-12, atol=1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("weighting", ["natural", "uniform", "robust"])
def test_psf_peak_is_one_with_weighting(weighting):
    """test_imaging.py:613-626: PSF peak == 1 for natural, uniform and robust weighting."""
    from rascil_amd import invert_blockvisibility
    from rascil_amd.weighting import weight_visibility

    bvis = _obs(nchan=2)
    model = create_image(128, 5e-4, bvis.phasecentre, [1.05e8], channel_bandwidth=[2e7])
    bvis = weight_visibility(bvis, model, weighting=weighting)
    psf, _ = invert_blockvisibility(bvis, model, dopsf=True, epsilon=1e-12)
    assert abs(psf.pixels.max() - 1.0) < 1e-10


@pytest.mark.gpu
def test_cellsize_too_large_raises():
    from rascil_amd.weighting import weight_visibility

    bvis = _obs()
    model = create_image(16, 1e-2, bvis.phasecentre, [1.05e8], channel_bandwidth=[2e7])
    with pytest.raises(ValueError):
        weight_visibility(bvis, model, weighting="uniform")


def _abi_weight(uvw, freq, wt, ntimes, vis_to_im, nimchan, nx, ny, cdelt0, cdelt1, mode, rob):
    """hipng_weight_visibility through the C-ABI on device copies; returns (weights, grid)."""
    import ctypes

    import torch

    from rascil_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda", 0)
    nrow, nchan, npol = wt.shape
    d2r = numpy.pi / 180.0
    u = torch.tensor(numpy.ascontiguousarray(uvw), device=dev)
    w = torch.tensor(numpy.ascontiguousarray(wt), device=dev)
    grid = torch.full((nimchan, ny, nx, npol), numpy.nan, dtype=torch.float64, device=dev)
    fq = numpy.ascontiguousarray(freq, "f8")
    v2i = numpy.ascontiguousarray(vis_to_im, numpy.int32)
    _lib.check(lib.hipng_weight_visibility(
        ctypes.c_void_p(u.data_ptr()), _lib.dptr(fq), ntimes, nrow // ntimes, nchan, npol,
        v2i.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), nimchan, nx, ny, 1.0 / (nx * cdelt0 * d2r),
        1.0 / (ny * cdelt1 * d2r), mode, float(rob), ctypes.c_void_p(w.data_ptr()),
        ctypes.c_void_p(grid.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
    torch.cuda.synchronize()
    return w.cpu().numpy(), grid.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("weighting,rob", [("uniform", 0.0), ("robust", 0.5)])
def test_gpu_weighting_production_tracks(weighting, rob):
    """Production-shape geometry (512 stations, 128 times over 4 h, 64 channels, 4 pols, 8192^2
    grid at 1.1e-5 rad) on 400 of its baselines over every time (short core baselines stay on
    one cell for several times, long ones move several cells per time), 10 % zero weights:
    weights and the density grid equal the oracle's."""
    from rascil_amd import simulation

    ntimes = 128
    uvw, freq, _ = simulation.config1_geometry(ntimes=ntimes)
    nbl = uvw.shape[0] // ntimes
    rng = numpy.random.default_rng(4)
    bl = numpy.sort(rng.choice(nbl, 400, replace=False))
    su = uvw.reshape(ntimes, nbl, 3)[:, bl].reshape(-1, 3)
    wt = rng.uniform(0.5, 1.5, (su.shape[0], len(freq), 4)) * (rng.uniform(size=(su.shape[0], len(freq), 4)) > 0.1)
    N, cell = 8192, 1.1e-5
    args = (numpy.zeros(len(freq), int), 1, N, N, numpy.degrees(-cell), numpy.degrees(cell))
    ref, rgrid = ow.weight(su, freq, wt, *args, weighting, robustness=rob)
    got, grid = _abi_weight(su, freq, wt, ntimes, *args, 1 if weighting == "uniform" else 2, rob)
    numpy.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)
    numpy.testing.assert_allclose(grid, rgrid.transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("nchan,npol", [(70, 2), (1, 1), (5, 4)])
def test_gpu_weighting_channel_blocks_and_pols(nchan, npol):
    """More channels than a wave holds (two channel blocks), one channel (64 baselines per
    wave), several image channels: weights and grid equal the oracle's."""
    rng = numpy.random.default_rng(nchan)
    ants = simulation.gaussian_disc_stations(30, 300.0, 1500.0, rng)
    ha = numpy.linspace(-2, 2, 37) * math.pi / 12
    freq = numpy.linspace(1e8, 1.6e8, nchan)
    from rascil_amd.simulation import baseline_uvw

    uvw = baseline_uvw(ants, ha, math.radians(-45), autocorrelations=False).reshape(-1, 3)
    wt = rng.uniform(0.1, 2.0, (uvw.shape[0], nchan, npol))
    nimchan = 3 if nchan >= 3 else 1
    v2i = numpy.minimum(numpy.arange(nchan) * nimchan // nchan, nimchan - 1)
    args = (v2i, nimchan, 256, 200, -0.01, 0.012)
    for mode, name in ((1, "uniform"), (2, "robust")):
        ref, rgrid = ow.weight(uvw, freq, wt, *args, name, robustness=-0.5)
        got, grid = _abi_weight(uvw, freq, wt, len(ha), *args, mode, -0.5)
        numpy.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-300)
        numpy.testing.assert_allclose(grid, rgrid.transpose(0, 2, 3, 1), rtol=1e-12, atol=1e-300)
