"""The FP64 image-side transform of grids of G = 8 M x 128 points, M in {10, 12, 16, 20, 24}
(8192^2 images at oversampling 1.25, 1.5, 2 and 2.5; a 16384^2 image at 1.5 for 24576):
invert and predict run it as a four-step transform (8 M x 128) fused with the w-screen
instead of transpose + rocFFT + screen. Checked through the C-ABI against the FP64 brute-force sums
(oracle): invert at probe pixels, predict of images that are zero off a set of pixels. The
pixels are the image edges, the centre, the band limits of the grid and random ones; w-stacking
off and on; one and two polarisations."""

import numpy
import pytest

from conftest import rel_l2
from oracle import explicit as ox
from rascil_amd import _lib, wgridder

pytestmark = pytest.mark.gpu

N, PIX, EPS = 8192, 1.1e-5, 1e-7
# (image side, oversampling, grid length): every M of the four-step, and both w-stacking modes
# at the production length 12288
GRIDS = [(8192, 1.25, 10240, True), (8192, 1.5, 12288, False), (8192, 1.5, 12288, True),
         (8192, 2.0, 16384, True), (8192, 2.5, 20480, True), (16384, 1.5, 24576, True)]
# a rectangular image: the x grid takes the four-step (12288), y is 3072 long
RECT = (8192, 2048, 1.5, 12288, 3072)


def _obs(seed, nrow=60):
    rng = numpy.random.default_rng(seed)
    uvw = numpy.empty((nrow, 3))
    uvw[:, :2] = rng.normal(0.0, 1500.0, (nrow, 2))
    uvw[:, 2] = rng.normal(0.0, 300.0, nrow)
    freq = numpy.array([1.2e9, 1.3e9])
    return rng, uvw, freq


def _probes(rng, n=N, m=None):
    """distinct pixels of an n x m image: the 8 x 8 combinations of edge and centre rows /
    columns, 400 random"""
    m = n if m is None else m
    ex, ey = numpy.meshgrid(numpy.array([0, 1, 2, n // 2 - 1, n // 2, n // 2 + 1, n - 2, n - 1]),
                            numpy.array([0, 1, 2, m // 2 - 1, m // 2, m // 2 + 1, m - 2, m - 1]), indexing="ij")
    lin = numpy.unique(numpy.concatenate([(ex * m + ey).ravel(), rng.integers(0, n * m, 400)]))
    return lin // m, lin % m


@pytest.mark.parametrize("n,sigma,g,wst", GRIDS)
def test_fourstep_invert_one_polarisation(n, sigma, g, wst):
    rng, uvw, freq = _obs(21)
    pix = PIX * N / n
    ms = rng.normal(size=(uvw.shape[0], 2)) + 1j * rng.normal(size=(uvw.shape[0], 2))
    wgt = rng.uniform(0.5, 1.5, ms.shape)
    plan = {}
    _lib.profile_enable(True)
    try:
        d = wgridder.ms2dirty(uvw, freq, ms, wgt, n, n, pix, pix, 0, 0, EPS, wst, precision="fp64",
                              plan_out=plan, plan_opts={"sigma": sigma})
        prof = _lib.profile_read()
    finally:
        _lib.profile_enable(False)
    assert plan["gx"] == g and plan["gy"] == g
    # the four-step kernels ran in every pass (one profile entry per pass), and the transposes
    # and screen of the rocFFT image pass did not
    assert prof["fourstep_1"][1] >= 1 and prof["fourstep_2"][1] == prof["fourstep_1"][1]
    assert "transpose" not in prof and "screen_acc" not in prof
    ix, iy = _probes(rng, n)
    ref = ox.explicit_ms2dirty_at(uvw, freq, ms, wgt, ix, iy, n, n, pix, pix, wst)
    err = rel_l2(d[ix, iy], ref)
    print(f"G={g} wst={wst} planes={plan['nplanes']}: invert {err:.2e} at {ix.size} probe pixels")
    assert err <= EPS


# four polarisations: the FP64 plane batch is too large for out-of-place band transforms, so
# both directions run with in-place ones
@pytest.mark.parametrize("nb", [2, 4])
def test_fourstep_invert_batched_polarisations(nb):
    torch = pytest.importorskip("torch")
    rng, uvw, freq = _obs(22, nrow=40)
    ms = rng.normal(size=(uvw.shape[0], 2, nb)) + 1j * rng.normal(size=(uvw.shape[0], 2, nb))
    wgt = rng.uniform(0.5, 1.5, ms.shape)
    dev = torch.device("cuda", 0)
    geom = wgridder.Geometry(torch.from_numpy(uvw).to(dev), freq, N, N, PIX, PIX, EPS, True, "fp64",
                             nbatch=nb, plan_opts={"sigma": 1.5})
    try:
        assert geom.plan["gx"] == 12288
        out = geom.ms2dirty(torch.from_numpy(ms).to(dev), torch.from_numpy(wgt).to(dev))
        ix, iy = _probes(rng)
        sel = out[:, torch.from_numpy(ix).to(dev), torch.from_numpy(iy).to(dev)].cpu().numpy()
        for b in range(nb):
            ref = ox.explicit_ms2dirty_at(uvw, freq, ms[..., b], wgt[..., b], ix, iy, N, N, PIX, PIX, True)
            err = rel_l2(sel[b], ref)
            print(f"polarisation {b}: invert {err:.2e}")
            assert err <= EPS
        # a second call on the kept geometry stores again (no stale accumulator)
        out2 = geom.ms2dirty(torch.from_numpy(ms).to(dev), torch.from_numpy(wgt).to(dev))
        assert torch.equal(out, out2)
    finally:
        geom.close()


@pytest.mark.parametrize("n,sigma,g,wst", GRIDS)
def test_fourstep_predict_one_polarisation(n, sigma, g, wst):
    rng, uvw, freq = _obs(23)
    pix = PIX * N / n
    ix, iy = _probes(rng, n)
    val = rng.normal(size=ix.size)
    dirty = numpy.zeros((n, n))
    dirty[ix, iy] = val
    wgt = rng.uniform(0.5, 1.5, (uvw.shape[0], 2))
    plan = {}
    _lib.profile_enable(True)
    try:
        v = wgridder.dirty2ms(uvw, freq, dirty, wgt, pix, pix, 0, 0, EPS, wst, precision="fp64",
                              plan_out=plan, plan_opts={"sigma": sigma})
        prof = _lib.profile_read()
    finally:
        _lib.profile_enable(False)
    assert plan["gx"] == g and plan["gy"] == g
    assert prof["fourstep_1"][1] >= 1 and prof["fourstep_2"][1] == prof["fourstep_1"][1]
    assert "screen_fill" not in prof
    ref = ox.explicit_dirty2ms_sparse(uvw, freq, ix, iy, val, n, n, pix, pix, wgt, wst)
    err = rel_l2(v, ref)
    print(f"G={g} wst={wst} planes={plan['nplanes']}: predict {err:.2e} from {ix.size} pixels")
    assert err <= EPS


@pytest.mark.parametrize("nb", [2, 4])
def test_fourstep_predict_batched_polarisations(nb):
    torch = pytest.importorskip("torch")
    rng, uvw, freq = _obs(24, nrow=40)
    ix, iy = _probes(rng)
    val = rng.normal(size=(nb, ix.size))
    model = numpy.zeros((nb, N, N))
    for b in range(nb):
        model[b, ix, iy] = val[b]
    dev = torch.device("cuda", 0)
    geom = wgridder.Geometry(torch.from_numpy(uvw).to(dev), freq, N, N, PIX, PIX, EPS, True, "fp64",
                             nbatch=nb, plan_opts={"sigma": 1.5})
    try:
        assert geom.plan["gx"] == 12288
        md = torch.from_numpy(model).to(dev)
        v = geom.dirty2ms(md).cpu().numpy()
        for b in range(nb):
            ref = ox.explicit_dirty2ms_sparse(uvw, freq, ix, iy, val[b], N, N, PIX, PIX, None, True)
            err = rel_l2(v[..., b], ref)
            print(f"polarisation {b}: predict {err:.2e}")
            assert err <= EPS
        # repeated calls on the kept geometry give the same visibilities
        assert numpy.array_equal(geom.dirty2ms(md).cpu().numpy(), v)
    finally:
        geom.close()


def test_fourstep_rectangular_image():
    nx, ny, sigma, gx, gy = RECT
    rng, uvw, freq = _obs(25)
    ms = rng.normal(size=(uvw.shape[0], 2)) + 1j * rng.normal(size=(uvw.shape[0], 2))
    wgt = rng.uniform(0.5, 1.5, ms.shape)
    ix, iy = _probes(rng, nx, ny)
    plan = {}
    _lib.profile_enable(True)
    try:
        d = wgridder.ms2dirty(uvw, freq, ms, wgt, nx, ny, PIX, PIX, 0, 0, EPS, True, precision="fp64",
                              plan_out=plan, plan_opts={"sigma": sigma})
        val = rng.normal(size=ix.size)
        dirty = numpy.zeros((nx, ny))
        dirty[ix, iy] = val
        v = wgridder.dirty2ms(uvw, freq, dirty, wgt, PIX, PIX, 0, 0, EPS, True, precision="fp64",
                              plan_opts={"sigma": sigma})
        prof = _lib.profile_read()
    finally:
        _lib.profile_enable(False)
    assert plan["gx"] == gx and plan["gy"] == gy
    assert prof["fourstep_1"][1] >= 2 and "screen_fill" not in prof and "screen_acc" not in prof
    ei = rel_l2(d[ix, iy], ox.explicit_ms2dirty_at(uvw, freq, ms, wgt, ix, iy, nx, ny, PIX, PIX, True))
    ep = rel_l2(v, ox.explicit_dirty2ms_sparse(uvw, freq, ix, iy, val, nx, ny, PIX, PIX, wgt, True))
    print(f"{nx} x {ny}: invert {ei:.2e}, predict {ep:.2e}")
    assert ei <= EPS and ep <= EPS


def test_fourstep_graph_replays_match_eager():
    """invert + predict of a kept FP64 geometry on a four-step length (G = 10240), captured in a
    hipGraph; every replay, with eager work in between, matches the eager result"""
    torch = pytest.importorskip("torch")
    rng, uvw, freq = _obs(26, nrow=30)
    dev = torch.device("cuda", 0)
    ms = torch.from_numpy(rng.normal(size=(uvw.shape[0], 2)) + 1j * rng.normal(size=(uvw.shape[0], 2))).to(dev)
    img = torch.from_numpy(rng.normal(size=(N, N))).to(dev)
    geom = wgridder.Geometry(torch.from_numpy(uvw).to(dev), freq, N, N, PIX, PIX, EPS, True, "fp64",
                             plan_opts={"sigma": 1.25})
    try:
        assert geom.plan["gx"] == 10240
        d_out = torch.empty((1, N, N), dtype=torch.float64, device=dev)
        v_out = torch.empty((uvw.shape[0], 2, 1), dtype=torch.complex128, device=dev)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):  # warm-up on the capture stream
            geom.ms2dirty(ms, out=d_out)
            geom.dirty2ms(img, out=v_out)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        d_ref, v_ref = d_out.clone(), v_out.clone()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            geom.ms2dirty(ms, out=d_out)
            geom.dirty2ms(img, out=v_out)
        for _ in range(3):
            d_out.fill_(7.0)
            v_out.fill_(7.0)
            torch.cuda.synchronize()
            graph.replay()
            torch.cuda.synchronize()
            # FP64 tiles flush by atomics: equal up to summation order
            assert (torch.linalg.norm(d_out - d_ref) / torch.linalg.norm(d_ref)).item() < 1e-12
            assert (torch.linalg.norm(v_out - v_ref) / torch.linalg.norm(v_ref)).item() < 1e-12
            _ = (ms * 2).sum().item()  # eager work between replays
        del graph
    finally:
        geom.close()
