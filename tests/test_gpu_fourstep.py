"""The FP64 image-side transform of 12288-point grids (an 8192^2 image at oversampling 1.5):
invert runs it as a four-step transform (96 x 128) fused with the w-screen instead of
transpose + rocFFT + screen. Checked through the C-ABI against the FP64 brute-force sums
(oracle) at probe pixels: the image edges, the centre, the band limits of the grid and
random pixels, with w-stacking off and on, for one and two polarisations."""

import numpy
import pytest

from conftest import rel_l2
from oracle import explicit as ox
from rascil_amd import _lib, wgridder

pytestmark = pytest.mark.gpu

N, PIX, EPS = 8192, 1.1e-5, 1e-7


def _obs(seed, nrow=60):
    rng = numpy.random.default_rng(seed)
    uvw = numpy.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0.0, 1500.0, (nrow, 2))
    uvw[:, 2] = rng.normal(0.0, 300.0, nrow)
    freq = numpy.array([1.2e9, 1.3e9])
    return rng, uvw, freq


def _probes(rng):
    edge = numpy.array([0, 1, 2, N // 2 - 1, N // 2, N // 2 + 1, N - 2, N - 1])
    ex, ey = numpy.meshgrid(edge, edge, indexing="ij")
    ix = numpy.concatenate([ex.ravel(), rng.integers(0, N, 400)])
    iy = numpy.concatenate([ey.ravel(), rng.integers(0, N, 400)])
    return ix, iy


@pytest.mark.parametrize("wst", [False, True])
def test_fourstep_invert_one_polarisation(wst):
    rng, uvw, freq = _obs(21)
    ms = rng.normal(size=(uvw.shape[0], 2)) + 1j * rng.normal(size=(uvw.shape[0], 2))
    wgt = rng.uniform(0.5, 1.5, ms.shape)
    plan = {}
    _lib.profile_enable(True)
    try:
        d = wgridder.ms2dirty(uvw, freq, ms, wgt, N, N, PIX, PIX, 0, 0, EPS, wst, precision="fp64",
                              plan_out=plan, plan_opts={"sigma": 1.5})
        prof = _lib.profile_read()
    finally:
        _lib.profile_enable(False)
    assert plan["gx"] == 12288 and plan["gy"] == 12288
    # the four-step kernels ran, one stage-2 launch per plane, and the rocFFT image pass did not
    assert prof["fourstep_1"][1] >= 1 and prof["fourstep_2"][1] == plan["nplanes"]
    assert "transpose" not in prof and "screen_acc" not in prof
    ix, iy = _probes(rng)
    ref = ox.explicit_ms2dirty_at(uvw, freq, ms, wgt, ix, iy, N, N, PIX, PIX, wst)
    err = rel_l2(d[ix, iy], ref)
    print(f"wst={wst} planes={plan['nplanes']}: invert {err:.2e} at {ix.size} probe pixels")
    assert err <= EPS


def test_fourstep_invert_two_polarisations():
    torch = pytest.importorskip("torch")
    rng, uvw, freq = _obs(22, nrow=40)
    ms = rng.normal(size=(uvw.shape[0], 2, 2)) + 1j * rng.normal(size=(uvw.shape[0], 2, 2))
    wgt = rng.uniform(0.5, 1.5, ms.shape)
    dev = torch.device("cuda", 0)
    geom = wgridder.Geometry(torch.from_numpy(uvw).to(dev), freq, N, N, PIX, PIX, EPS, True, "fp64",
                             nbatch=2, plan_opts={"sigma": 1.5})
    try:
        assert geom.plan["gx"] == 12288
        out = geom.ms2dirty(torch.from_numpy(ms).to(dev), torch.from_numpy(wgt).to(dev))
        ix, iy = _probes(rng)
        sel = out[:, torch.from_numpy(ix).to(dev), torch.from_numpy(iy).to(dev)].cpu().numpy()
        for b in range(2):
            ref = ox.explicit_ms2dirty_at(uvw, freq, ms[..., b], wgt[..., b], ix, iy, N, N, PIX, PIX, True)
            err = rel_l2(sel[b], ref)
            print(f"polarisation {b}: invert {err:.2e}")
            assert err <= EPS
        # a second call on the kept geometry stores again (no stale accumulator)
        out2 = geom.ms2dirty(torch.from_numpy(ms).to(dev), torch.from_numpy(wgt).to(dev))
        assert torch.equal(out, out2)
    finally:
        geom.close()
