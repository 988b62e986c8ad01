"""Gain application: the oracle pinned by known answers and closed forms (CPU), the GPU
kernel (through hipng_apply_gaintable) against the oracle, and the reference's round-trip
property (tests/processing_components/test_calibration_operations.py:78-181)."""

import math

import numpy
import pytest

from oracle import gains as og
from rascil_amd import Direction, PolarisationFrame, simulation


def _one(vis4, g1, g2, npol=4, inverse=False):
    """One time, one baseline (antennas 0, 1), one channel."""
    vis = numpy.asarray(vis4, complex).reshape(1, 1, 1, npol)
    nrec = 1 if npol == 1 else 2
    gain = numpy.stack([numpy.asarray(g1, complex).reshape(nrec, nrec),
                        numpy.asarray(g2, complex).reshape(nrec, nrec)])[None, :, None]
    gw = numpy.full(gain.shape, 1.0)
    gw[0, 0] = 2.0
    gw[0, 1] = 3.0
    return og.apply_gaintable(vis, numpy.ones(vis.shape), numpy.zeros(vis.shape), numpy.zeros(1), [0], [1],
                              gain, gw, numpy.zeros(1), numpy.ones(1), inverse=inverse)


def test_oracle_known_answer_conj_without_transpose():
    """G1 V conj(G2) by hand: G1 = [[1, 2i], [0, 1]], V = I, G2 = [[1, 0], [3i, 1]] gives
    [[7, 2i], [-3i, 1]] (G1 V G2^H would give [[1, -i], [0, 1]])."""
    v, w = _one([1, 0, 0, 1], [1, 2j, 0, 1], [1, 0, 3j, 1])
    numpy.testing.assert_allclose(v.ravel(), [7, 2j, -3j, 1], rtol=0, atol=1e-15)
    numpy.testing.assert_array_equal(w.ravel(), [1, 1, 1, 1])


def test_oracle_known_answer_diagonal_two_pols():
    """npol 2: diag(G1 diag(v) conj(G2)); with G1 = [[1, 2], [0, 1]], G2 = [[1, 0], [1j, 1]],
    v = (1, 1): [[1, 2], [0, 1]] diag(1, 1) [[1, 0], [-1j, 1]] = [[1 - 2i, 2], [-1i, 1]]."""
    v, _ = _one([1, 1], [1, 2, 0, 1], [1, 0, 1j, 1], npol=2)
    numpy.testing.assert_allclose(v.ravel(), [1 - 2j, 1], rtol=0, atol=1e-15)


def test_oracle_scalar_phase_and_weight():
    """npol 1, g = e^{i phi}: V e^{i (phi1 - phi2)}, weight = gainwt1 * gainwt2."""
    v, w = _one([2 + 1j], [numpy.exp(0.3j)], [numpy.exp(-0.5j)], npol=1)
    numpy.testing.assert_allclose(v.ravel(), [(2 + 1j) * numpy.exp(0.8j)], rtol=1e-15)
    numpy.testing.assert_array_equal(w.ravel(), [6.0])
    v, w = _one([2 + 1j], [0.0], [1.0], npol=1)
    assert v.ravel()[0] == 0 and w.ravel()[0] == 0
    v, w = _one([2 + 1j], [0.0], [1.0], npol=1, inverse=True)  # 1/0 -> 0
    assert v.ravel()[0] == 0 and w.ravel()[0] == 0


def test_oracle_singular_inverse():
    v, w = _one([1, 2, 3, 4], [1, 1, 1, 1], [1, 0, 0, 1], inverse=True)
    assert numpy.all(v == 0) and numpy.all(w == 0)
    v, w = _one([5, 6], [1, 1, 1, 1], [1, 0, 0, 1], npol=2, inverse=True)
    numpy.testing.assert_array_equal(v.ravel(), [0, 6])
    numpy.testing.assert_array_equal(w.ravel(), [0, 1])


def _obs(pol, nchan=3, ntimes=5, seed=0):
    rng = numpy.random.default_rng(seed)
    ants = simulation.gaussian_disc_stations(12, 200.0, 800.0, rng)
    ha = numpy.linspace(-1, 1, ntimes) * math.pi / 12
    bvis = simulation.create_blockvisibility(ants, ha, numpy.linspace(1e8, 1.1e8, nchan),
                                             Direction(0.0, math.radians(-35)), PolarisationFrame(pol))
    bvis.vis = rng.normal(size=bvis.vis.shape) + 1j * rng.normal(size=bvis.vis.shape)
    bvis.weight = rng.uniform(0.5, 2.0, bvis.vis.shape)
    return bvis


def _random_gains(gt, rng, amp=0.1, phase=0.5, offdiag=0.0):
    g = gt.gain
    shape = g[..., 0, 0].shape
    for k in range(g.shape[-1]):
        g[..., k, k] = (1 + amp * rng.normal(size=shape)) * numpy.exp(1j * phase * rng.normal(size=shape))
    if g.shape[-1] == 2 and offdiag:
        g[..., 0, 1] = offdiag * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
        g[..., 1, 0] = offdiag * (rng.normal(size=shape) + 1j * rng.normal(size=shape))
    gt.weight = rng.uniform(0.5, 1.5, g.shape)
    return gt


def _oracle(bvis, gt, inverse=False, use_flags=False):
    return og.apply_gaintable(bvis.vis, bvis.weight, bvis.flags, bvis.time, bvis.baselines[0], bvis.baselines[1],
                              gt.gain, gt.weight, gt.time, gt.interval, inverse=inverse, use_flags=use_flags)


@pytest.mark.parametrize("pol", ["stokesI", "linear", "circular"])
def test_oracle_round_trip(pol):
    """apply then apply inverse restores the visibilities (the reference's own property)."""
    from rascil_amd.calibration import create_gaintable_from_blockvisibility

    for jones in ("T", "G", "B"):
        bvis = _obs(pol)
        gt = _random_gains(create_gaintable_from_blockvisibility(bvis, jones_type=jones),
                           numpy.random.default_rng(1), offdiag=0.05)
        v1, w1 = _oracle(bvis, gt)
        bvis2 = _obs(pol)
        bvis2.vis, bvis2.weight = v1, w1
        v2, _ = _oracle(bvis2, gt, inverse=True)
        assert numpy.max(numpy.abs(v2 - bvis.vis)) < 1e-12


def test_oracle_table_channels_and_repeated_rows():
    """A one-channel (T) table changes channel 0 only; a time matched by two rows gets both
    gains in row order."""
    from rascil_amd.calibration import GainTable

    bvis = _obs("stokesI", ntimes=1)
    nants = 12
    g = numpy.exp(1j * numpy.arange(nants))[None, :, None, None, None] * numpy.ones((2, 1, 1, 1, 1))
    g[1] *= 2.0
    gt = GainTable(g, [bvis.time[0], bvis.time[0]], [1.0, 1.0])
    v, _ = _oracle(bvis, gt)
    a1, a2 = bvis.baselines
    s = numpy.exp(1j * (a1 - a2)) * numpy.exp(1j * (a1 - a2)) * 4.0
    numpy.testing.assert_allclose(v[0, :, 0, 0], bvis.vis[0, :, 0, 0] * s, rtol=1e-14)
    numpy.testing.assert_array_equal(v[:, :, 1:], bvis.vis[:, :, 1:])


def _gpu_case(pol, jones, inverse, use_flags, offdiag, timeslice=None, seed=0, widen=1.0):
    from rascil_amd.calibration import apply_gaintable, create_gaintable_from_blockvisibility

    bvis = _obs(pol, seed=seed)
    rng = numpy.random.default_rng(seed + 7)
    if use_flags:
        bvis.flags = (rng.uniform(size=bvis.vis.shape) < 0.2).astype(int)
    gt = _random_gains(create_gaintable_from_blockvisibility(bvis, timeslice=timeslice, jones_type=jones), rng,
                       offdiag=offdiag)
    gt.interval = gt.interval * widen
    if inverse and gt.gain.shape[-1] == 2:
        gt.gain[0, 3, 0] = [[1, 2], [2, 4]]  # singular for antenna 3 at the first row
    if inverse and gt.gain.shape[-1] == 1:
        gt.gain[0, 3, 0] = 0.0
    ref_v, ref_w = _oracle(bvis, gt, inverse=inverse, use_flags=use_flags)
    out = apply_gaintable(bvis, gt, inverse=inverse, use_flags=use_flags)
    assert out is bvis
    return out, ref_v, ref_w


@pytest.mark.gpu
@pytest.mark.parametrize("pol", ["stokesI", "linear", "circular", "linearnp"])
@pytest.mark.parametrize("jones", ["T", "B"])
@pytest.mark.parametrize("inverse", [False, True])
def test_gpu_apply_matches_oracle(pol, jones, inverse):
    out, ref_v, ref_w = _gpu_case(pol, jones, inverse, False, 0.1)
    numpy.testing.assert_allclose(out.vis, ref_v, rtol=1e-12, atol=1e-13)
    numpy.testing.assert_allclose(out.weight, ref_w, rtol=1e-14, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("pol", ["stokesI", "linear", "linearnp"])
def test_gpu_apply_use_flags_and_time_bins(pol):
    """use_flags masks the flagged row blocks; binned gain rows (timeslice) whose intervals
    overlap neighbouring times apply several rows to one time, in row order."""
    out, ref_v, ref_w = _gpu_case(pol, "G", False, True, 0.05, timeslice=0.05, widen=2.5)
    numpy.testing.assert_allclose(out.vis, ref_v, rtol=1e-12, atol=1e-13)
    numpy.testing.assert_allclose(out.weight, ref_w, rtol=1e-14, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("pol", ["stokesI", "linear", "circular"])
@pytest.mark.parametrize("jones", ["T", "G", "B"])
def test_gpu_round_trip(pol, jones):
    """test_calibration_operations.py:162-181: apply then apply the inverse: < 1e-12."""
    from rascil_amd.calibration import apply_gaintable, create_gaintable_from_blockvisibility

    bvis = _obs(pol)
    original = bvis.vis.copy()
    gt = _random_gains(create_gaintable_from_blockvisibility(bvis, jones_type=jones), numpy.random.default_rng(3),
                       amp=0.1, phase=0.1, offdiag=0.05)
    apply_gaintable(bvis, gt)
    assert numpy.max(numpy.abs(bvis.vis - original)) > 0.0  # :123-139
    apply_gaintable(bvis, gt, inverse=True)
    assert numpy.max(numpy.abs(bvis.vis - original)) < 1e-12
