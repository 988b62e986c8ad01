"""Run-to-run behaviour of the GPU path. Predict (dirty2ms) gathers without atomics and is
bit-identical between runs in every precision, through one-shot calls and through a kept
geometry. Invert (ms2dirty) adds grid cells with float atomics in an order that varies
between runs, so its images agree between runs only to the rounding of that order (FP64
1e-14, FP32 modes 1e-6 relative) and every run meets epsilon against the brute-force sum."""

import numpy
import pytest

from conftest import rel_l2
from oracle import explicit as ox
from rascil_amd import wgridder

pytestmark = pytest.mark.gpu

FREQ = numpy.array([1.0e8, 1.13e8, 1.31e8])
RUN_TOL = {"fp64": 1e-14, "mixed": 1e-6, "fp32": 1e-6}
EPS = {"fp64": 1e-10, "mixed": 1e-5, "fp32": 1e-3}


def _inputs(nrow, nb, seed):
    rng = numpy.random.default_rng(seed)
    shape = (nrow, len(FREQ), nb)
    ms = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    wgt = rng.uniform(0.2, 1.5, shape)
    dirty = rng.normal(size=(nb, 96, 80))
    return ms, wgt, dirty


@pytest.mark.parametrize("prec", ["fp64", "mixed", "fp32"])
@pytest.mark.parametrize("nb", [1, 4])
def test_predict_repeats_bit_for_bit_and_invert_to_rounding(small_obs, prec, nb):
    uvw = small_obs[0]
    ms, wgt, dirty = _inputs(uvw.shape[0], nb, seed=nb)
    eps = EPS[prec]
    args = (96, 80, 6e-4, 7e-4)
    vis = [wgridder.dirty2ms(uvw, FREQ, dirty, wgt, 6e-4, 7e-4, epsilon=eps, do_wstacking=True, precision=prec)
           for _ in range(3)]
    for v in vis[1:]:
        assert numpy.array_equal(v, vis[0])
    img = [wgridder.ms2dirty(uvw, FREQ, ms, wgt, *args, epsilon=eps, do_wstacking=True, precision=prec)
           for _ in range(3)]
    for im in img[1:]:
        assert rel_l2(im, img[0]) < RUN_TOL[prec]
    for b in range(nb):
        ref = ox.explicit_ms2dirty(uvw, FREQ, ms[..., b], wgt[..., b], *args, True)
        assert rel_l2(img[0][b], ref) < eps
        refv = ox.explicit_dirty2ms(uvw, FREQ, dirty[b], wgt[..., b], 6e-4, 7e-4, True)
        assert rel_l2(vis[0][..., b], refv) < eps


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_kept_geometry_repeats(small_obs, prec):
    """A kept geometry gives the one-shot call's predict bit for bit on every execution, and
    inverts that agree with each other to the rounding of the accumulation order."""
    import torch

    uvw = small_obs[0]
    ms, wgt, dirty = _inputs(uvw.shape[0], 2, seed=7)
    eps = EPS[prec]
    dev = torch.device("cuda", 0)
    cd = torch.complex128 if prec == "fp64" else torch.complex64
    rd = torch.float64 if prec == "fp64" else torch.float32
    one = wgridder.dirty2ms(uvw, FREQ, dirty, wgt, 6e-4, 7e-4, epsilon=eps, do_wstacking=True, precision=prec)
    g = wgridder.Geometry(uvw, FREQ, 96, 80, 6e-4, 7e-4, eps, True, prec, nbatch=2)
    try:
        td = torch.tensor(dirty, dtype=rd, device=dev)
        tw = torch.tensor(wgt, dtype=rd, device=dev)
        tm = torch.tensor(ms, dtype=cd, device=dev)
        vis = [g.dirty2ms(td, tw).cpu().numpy() for _ in range(3)]
        img = [g.ms2dirty(tm, tw).cpu().numpy() for _ in range(3)]
    finally:
        g.close()
    for v in vis:
        assert numpy.array_equal(v, one)
    for im in img[1:]:
        assert rel_l2(im, img[0]) < RUN_TOL[prec]
