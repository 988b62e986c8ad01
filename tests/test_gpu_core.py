"""GPU parity of the hipng kernels through the C-ABI: bit-exact bucket sort against the
CPU oracle sort, and ms2dirty / dirty2ms against the FP64 brute-force sums (oracle)."""

import numpy
import pytest

from conftest import rel_l2, rel_max
from oracle import explicit as ox
from oracle import sortref
from rascil_amd import simulation, wgridder

pytestmark = pytest.mark.gpu

C = 299792458.0
FREQ2 = numpy.array([1.0e8, 1.31e8])


def _case(uvw, nchan_freq, seed=0, nb=None):
    rng = numpy.random.default_rng(seed)
    nrow = uvw.shape[0]
    nch = len(nchan_freq)
    shape = (nrow, nch) if nb is None else (nrow, nch, nb)
    ms = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    wgt = rng.uniform(0.2, 1.5, shape)
    return ms, wgt


def _extent_plan(uvw, freq, nx, ny, px, py, eps, wst, prec, plan_opts=None):
    e = wgridder.uvw_extent(uvw, freq)
    return wgridder.plan_params(e[0], e[1], e[2], e[3], nx, ny, px, py, eps, wst, prec, plan_opts=plan_opts)


@pytest.mark.parametrize("wst", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nchan", [1, 16])
def test_bucket_sort_bit_exact(small_obs, wst, prec, nchan):
    uvw, bvis = small_obs
    uvw = bvis.uvw.reshape(-1, 3)  # full C0 geometry: 148,500 rows
    freq = numpy.linspace(1e8, 1.3e8, nchan)
    plan = _extent_plan(uvw, freq, 512, 512, 5e-4, 5e-4, 1e-5 if prec == "fp64" else 1e-4, wst, prec)
    keys, perm = wgridder.bucket_sort(uvw, freq, plan)
    sx, sy, sw = wgridder.channel_scales(plan, freq)
    rk, rp = sortref.bucket_sort(uvw, sx, sy, sw, plan.as_dict())
    assert numpy.array_equal(keys, rk)
    assert numpy.array_equal(perm, rp)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("px", [2e-3, 8e-3])
def test_bucket_sort_bit_exact_wrapping_footprints(small_obs, prec, px):
    """An undersampled 64^2 image (cell 2e-3 / 8e-3 rad against a longest baseline of ~900
    wavelengths): grid coordinates leave the grid and wrap once (2e-3) or several times
    (8e-3), the path where the cell index takes a remainder instead of one add or subtract."""
    uvw, _ = small_obs
    freq = numpy.array([1.0e8, 1.15e8, 1.3e8])
    plan = _extent_plan(uvw, freq, 64, 64, px, px, 1e-5 if prec == "fp64" else 1e-3, True, prec)
    assert plan.umax * px > (0.6 if px < 5e-3 else 3.0)  # how far footprints reach past the grid
    keys, perm = wgridder.bucket_sort(uvw, freq, plan)
    sx, sy, sw = wgridder.channel_scales(plan, freq)
    rk, rp = sortref.bucket_sort(uvw, sx, sy, sw, plan.as_dict())
    assert numpy.array_equal(keys, rk)
    assert numpy.array_equal(perm, rp)


def test_wrapping_footprints_match_explicit_sums(small_obs):
    """The same several-times-wrapping 64^2 image through invert and predict: a periodic
    grid makes the wrapped coordinates exact, so both meet epsilon against brute force."""
    uvw, _ = small_obs
    freq = numpy.array([1.0e8, 1.3e8])
    ms, wgt = _case(uvw, freq, seed=3)
    px, eps = 8e-3, 1e-9
    d = wgridder.ms2dirty(uvw, freq, ms, wgt, 64, 64, px, px, 0, 0, eps, True)
    ref = ox.explicit_ms2dirty(uvw, freq, ms, wgt, 64, 64, px, px, True)
    assert rel_l2(d, ref) < 10 * eps
    model = numpy.random.default_rng(4).normal(size=(64, 64))
    v = wgridder.dirty2ms(uvw, freq, model, wgt, px, px, 0, 0, eps, True)
    vref = ox.explicit_dirty2ms(uvw, freq, model, wgt, px, px, True)
    assert rel_l2(v, vref) < 10 * eps


@pytest.mark.parametrize("pack,tile", [(True, 16), (False, 16), (True, 32)])
def test_bucket_sort_bit_exact_production_subsample(pack, tile):
    """1 % of the 512-station / 64-channel production geometry, 8192^2 image plan; with the
    bucket key carried through the digit passes in the sort words (pack 1, the default when
    the bits fit) or recomputed from uvw after them (pack 0), and with 32 x 32 tiles (two
    digit passes, the second also sees low key bits)."""
    opts = {"sort_pack": pack, "tile": tile}
    uvw, freq, _ = simulation.config1_geometry(ntimes=128, nchan=64)
    rng = numpy.random.default_rng(0)
    rows = numpy.sort(rng.choice(uvw.shape[0], uvw.shape[0] // 100, replace=False))
    uvw = uvw[rows]
    plan = _extent_plan(uvw, freq, 8192, 8192, 1.1e-5, 1.1e-5, 1e-5, True, "mixed", opts)
    assert plan.tile == tile
    keys, perm = wgridder.bucket_sort(uvw, freq, plan, plan_opts=opts)
    sx, sy, sw = wgridder.channel_scales(plan, freq)
    rk, rp = sortref.bucket_sort(uvw, sx, sy, sw, plan.as_dict())
    assert numpy.array_equal(keys, rk)
    assert numpy.array_equal(perm, rp)


CASES = [  # precision, epsilon, tolerance on the relative L2 error
    ("fp64", 1e-12, 1e-11),
    ("fp64", 1e-7, 1e-7),
    ("fp64", 1e-4, 1e-4),
    ("mixed", 1e-5, 1e-5),
    ("mixed", 1e-3, 1e-3),
    ("fp32", 1e-3, 1e-3),
]


@pytest.mark.parametrize("prec,eps,tol", CASES)
@pytest.mark.parametrize("wst", [False, True])
def test_ms2dirty_vs_explicit(small_obs, prec, eps, tol, wst):
    uvw, _ = small_obs
    nx, ny, px, py = 96, 80, 6e-4, 7e-4
    ms, wgt = _case(uvw, FREQ2)
    if prec != "fp64":
        ms, wgt = ms.astype("c8"), wgt.astype("f4")
    plan = {}
    d = wgridder.ms2dirty(uvw, FREQ2, ms, wgt, nx, ny, px, py, 0, 0, eps, wst,
                          precision=prec, plan_out=plan)
    ref = ox.explicit_ms2dirty(uvw, FREQ2, ms.astype("c16"), wgt.astype("f8"), nx, ny, px, py, wst)
    err = rel_l2(d, ref)
    print(f"ms2dirty {prec} eps={eps} wst={wst}: rel_l2={err:.3e} rel_max={rel_max(d, ref):.3e} "
          f"W={plan['W']} sigma={plan['sigma']} planes={plan['nplanes']}")
    assert err <= tol


@pytest.mark.parametrize("prec,eps,tol", [("mixed", 1e-5, 1e-5), ("fp32", 1e-3, 1e-3), ("fp64", 1e-7, 1e-7)])
def test_odd_half_height_image(small_obs, prec, eps, tol):
    """An image whose half-height is odd (ny = 90): the image columns around the grid wrap
    (the fftshift of an odd half) in every precision, both directions."""
    uvw, _ = small_obs
    nx, ny, px, py = 64, 90, 6e-4, 7e-4
    ms, wgt = _case(uvw, FREQ2)
    if prec != "fp64":
        ms, wgt = ms.astype("c8"), wgt.astype("f4")
    d = wgridder.ms2dirty(uvw, FREQ2, ms, wgt, nx, ny, px, py, 0, 0, eps, True, precision=prec)
    ref = ox.explicit_ms2dirty(uvw, FREQ2, ms.astype("c16"), wgt.astype("f8"), nx, ny, px, py, True)
    assert rel_l2(d, ref) <= tol
    rng = numpy.random.default_rng(11)
    img = rng.normal(size=(nx, ny)).astype("f8" if prec == "fp64" else "f4")
    v = wgridder.dirty2ms(uvw, FREQ2, img, None, px, py, 0, 0, eps, True, precision=prec)
    vref = ox.explicit_dirty2ms(uvw, FREQ2, img.astype("f8"), None, px, py, True)
    assert rel_l2(v, vref) <= tol


@pytest.mark.parametrize("prec,eps,tol", CASES)
@pytest.mark.parametrize("wst", [False, True])
def test_dirty2ms_vs_explicit(small_obs, prec, eps, tol, wst):
    uvw, _ = small_obs
    nx, ny, px, py = 96, 80, 6e-4, 7e-4
    rng = numpy.random.default_rng(7)
    dirty = rng.normal(size=(nx, ny))
    wgt = rng.uniform(0.2, 1.5, (uvw.shape[0], 2))
    if prec != "fp64":
        dirty, wgt = dirty.astype("f4"), wgt.astype("f4")
    plan = {}
    ms = wgridder.dirty2ms(uvw, FREQ2, dirty, wgt, px, py, 0, 0, eps, wst, precision=prec,
                           plan_out=plan)
    ref = ox.explicit_dirty2ms(uvw, FREQ2, dirty.astype("f8"), wgt.astype("f8"), px, py, wst)
    err = rel_l2(ms, ref)
    print(f"dirty2ms {prec} eps={eps} wst={wst}: rel_l2={err:.3e} rel_max={rel_max(ms, ref):.3e} "
          f"W={plan['W']} sigma={plan['sigma']} planes={plan['nplanes']}")
    assert err <= tol


def test_batch_equals_single(small_obs):
    uvw, _ = small_obs
    ms, wgt = _case(uvw, FREQ2, nb=3)
    d3 = wgridder.ms2dirty(uvw, FREQ2, ms, wgt, 64, 64, 6e-4, 6e-4, 0, 0, 1e-10, True)
    for b in range(3):
        d1 = wgridder.ms2dirty(uvw, FREQ2, numpy.ascontiguousarray(ms[..., b]),
                               numpy.ascontiguousarray(wgt[..., b]), 64, 64, 6e-4, 6e-4, 0, 0, 1e-10, True)
        assert rel_l2(d3[b], d1) < 1e-13
    img = numpy.random.default_rng(3).normal(size=(3, 64, 64))
    v3 = wgridder.dirty2ms(uvw, FREQ2, img, None, 6e-4, 6e-4, 0, 0, 1e-10, True)
    for b in range(3):
        v1 = wgridder.dirty2ms(uvw, FREQ2, img[b], None, 6e-4, 6e-4, 0, 0, 1e-10, True)
        assert rel_l2(v3[..., b], v1) < 1e-13


@pytest.mark.parametrize("prec,eps", [("fp64", 1e-12), ("mixed", 1e-5)])
def test_adjoint_identity_full_c0(small_obs, prec, eps):
    """<dirty2ms(I), V>_R = <I, ms2dirty(V)> on the full 148,500-visibility C0 set at 512^2."""
    _, bvis = small_obs
    uvw = bvis.uvw.reshape(-1, 3)
    rng = numpy.random.default_rng(11)
    V = rng.normal(size=(uvw.shape[0], 1)) + 1j * rng.normal(size=(uvw.shape[0], 1))
    I = rng.normal(size=(512, 512))
    cd, rd = ("c16", "f8") if prec == "fp64" else ("c8", "f4")
    a = wgridder.dirty2ms(uvw, [1e8], I.astype(rd), None, 5e-4, 5e-4, 0, 0, eps, True, precision=prec)
    b = wgridder.ms2dirty(uvw, [1e8], V.astype(cd), None, 512, 512, 5e-4, 5e-4, 0, 0, eps, True,
                          precision=prec)
    lhs = numpy.vdot(a.astype("c16"), V).real
    rhs = numpy.sum(I * b.astype("f8"))
    scale = numpy.linalg.norm(a) * numpy.linalg.norm(V)
    assert abs(lhs - rhs) / scale < 10 * eps


def test_hermitian_flip_gpu(small_obs):
    uvw, _ = small_obs
    ms, _ = _case(uvw, FREQ2)
    a = wgridder.ms2dirty(uvw, FREQ2, ms, None, 64, 64, 6e-4, 6e-4, 0, 0, 1e-12, True)
    b = wgridder.ms2dirty(-uvw, FREQ2, ms.conj(), None, 64, 64, 6e-4, 6e-4, 0, 0, 1e-12, True)
    assert rel_l2(a, b) < 1e-12


def test_no_rows_gives_zero_image():
    d = wgridder.ms2dirty(numpy.zeros((0, 3)), [1e8], numpy.zeros((0, 1), "c16"), None, 32, 32,
                          1e-3, 1e-3, 0, 0, 1e-7, True)
    assert d.shape == (32, 32) and not d.any()


def test_invalid_arguments_raise():
    uvw = numpy.random.default_rng(0).normal(0, 100, (10, 3))
    with pytest.raises(ValueError):
        wgridder.ms2dirty(uvw, [1e8], numpy.ones((10, 1), "c16"), None, 33, 32, 1e-3, 1e-3, 0, 0, 1e-7, True)
    with pytest.raises(ValueError):
        wgridder.dirty2ms(uvw, [1e8], numpy.ones((32, 32), "f4"), None, 1e-3, 1e-3, 0, 0, 1e-9, True,
                          precision="mixed")


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_geometry_reuse_matches_one_shot(small_obs, prec):
    import torch

    uvw, _ = small_obs
    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(21)
    cd, rd = (torch.complex128, torch.float64) if prec == "fp64" else (torch.complex64, torch.float32)
    ms = torch.tensor(rng.normal(size=(uvw.shape[0], 2, 2)) + 1j * rng.normal(size=(uvw.shape[0], 2, 2)),
                      dtype=cd, device=dev)
    img = torch.tensor(rng.normal(size=(2, 64, 48)), dtype=rd, device=dev)
    eps = 1e-10 if prec == "fp64" else 1e-5
    geom = wgridder.Geometry(uvw, FREQ2, 64, 48, 6e-4, 7e-4, eps, True, prec, nbatch=2)
    for _ in range(2):  # executions are repeatable
        d = geom.ms2dirty(ms)
        v = geom.dirty2ms(img)
    d1 = wgridder.ms2dirty(torch.tensor(uvw, device=dev), FREQ2, ms, None, 64, 48, 6e-4, 7e-4, 0, 0, eps, True,
                           precision=prec)
    v1 = wgridder.dirty2ms(torch.tensor(uvw, device=dev), FREQ2, img, None, 6e-4, 7e-4, 0, 0, eps, True,
                           precision=prec)
    tol = 1e-12 if prec == "fp64" else 1e-6
    assert (torch.linalg.norm(d - d1) / torch.linalg.norm(d1)).item() < tol
    assert (torch.linalg.norm(v - v1) / torch.linalg.norm(v1)).item() < tol
    geom.close()


def test_graph_capture_replays_match_eager(small_obs):
    """invert + predict captured in a hipGraph; every replay (with eager work in between)
    matches the eager result."""
    import torch

    uvw, _ = small_obs
    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(22)
    ms = torch.tensor(rng.normal(size=(uvw.shape[0], 2)) + 1j * rng.normal(size=(uvw.shape[0], 2)),
                      dtype=torch.complex64, device=dev)
    img = torch.tensor(rng.normal(size=(64, 64)), dtype=torch.float32, device=dev)
    geom = wgridder.Geometry(uvw, FREQ2, 64, 64, 6e-4, 6e-4, 1e-5, True, "mixed")
    d_out = torch.empty((1, 64, 64), dtype=torch.float32, device=dev)
    v_out = torch.empty((uvw.shape[0], 2, 1), dtype=torch.complex64, device=dev)
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
    for it in range(3):
        d_out.fill_(7.0)
        v_out.fill_(7.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(d_out, d_ref) or (torch.linalg.norm(d_out - d_ref) / torch.linalg.norm(d_ref)).item() < 1e-6
        assert (torch.linalg.norm(v_out - v_ref) / torch.linalg.norm(v_ref)).item() < 1e-6
        _ = (ms * 2).sum().item()  # eager work between replays
    geom.close()


def _dense_core(seed=31):
    """A compact uv core observed repeatedly (60 short baselines x 20 near-identical times x
    32 channels 1 kHz apart: hundreds of visibilities share a first cell and start plane)
    plus 600 scattered rows, so both gridding paths carry work."""
    rng = numpy.random.default_rng(seed)
    base = numpy.column_stack([rng.uniform(-40, 40, 60), rng.uniform(-40, 40, 60), rng.uniform(-15, 15, 60)])
    core = numpy.repeat(base, 20, axis=0) + rng.normal(0, 1e-3, (1200, 3))
    spread = numpy.column_stack([rng.uniform(-300, 300, 600), rng.uniform(-300, 300, 600),
                                 rng.uniform(-60, 60, 600)])
    uvw = numpy.concatenate([core, spread])
    freq = 1.0e8 + 1.0e3 * numpy.arange(32)
    return uvw, freq


@pytest.mark.parametrize("prec,eps", [("fp64", 1e-5), ("mixed", 1e-5), ("fp32", 1e-4)])
def test_dense_runs_match_explicit(prec, eps):
    """Runs of visibilities with a common first cell go through the dense-run gridder; with
    the scattered rest through the sub-tile gatherer, invert and predict still equal the
    brute-force sums to epsilon. Oversampling 2 gives support 7 (lanes outside the
    footprint idle) and more than half of the visibilities in dense runs."""
    import torch

    uvw, freq = _dense_core()
    rng = numpy.random.default_rng(32)
    nx, ny, px = 48, 48, 2e-3
    ms = rng.normal(size=(uvw.shape[0], freq.size)) + 1j * rng.normal(size=(uvw.shape[0], freq.size))
    img = rng.normal(size=(nx, ny))
    dev = torch.device("cuda", 0)
    cd, rd = (torch.complex128, torch.float64) if prec == "fp64" else (torch.complex64, torch.float32)
    geom = wgridder.Geometry(uvw, freq, nx, ny, px, px, eps, True, prec, plan_opts={"sigma": 2.0})
    st = geom.stats()
    assert st["dense_enabled"] and st["dense_tasks"] > 0
    assert 0.5 * st["nvis"] < st["dense_vis"] < st["nvis"]  # both paths in use
    d = geom.ms2dirty(torch.tensor(ms, dtype=cd, device=dev))[0].double().cpu().numpy()
    v = geom.dirty2ms(torch.tensor(img, dtype=rd, device=dev))[..., 0].cpu().numpy()
    geom.close()
    ref_d = ox.explicit_ms2dirty(uvw, freq, ms, None, nx, ny, px, px, True)
    ref_v = ox.explicit_dirty2ms(uvw, freq, img, None, px, px, True)
    assert rel_l2(d, ref_d) < eps
    assert rel_l2(v.astype(numpy.complex128), ref_v) < eps


@pytest.mark.parametrize("nb", [2, 3, 4])
def test_dense_runs_polarisations_match_explicit(nb):
    """The batched instances of both gridding paths and of both degridders (4 polarisations
    is the production shape's batch) against the brute-force sums, polarisation by
    polarisation."""
    import torch

    uvw, freq = _dense_core(seed=33)
    rng = numpy.random.default_rng(34)
    nx, ny, px, eps = 48, 40, 2e-3, 1e-5
    ms = rng.normal(size=(uvw.shape[0], freq.size, nb)) + 1j * rng.normal(size=(uvw.shape[0], freq.size, nb))
    img = rng.normal(size=(nb, nx, ny))
    dev = torch.device("cuda", 0)
    geom = wgridder.Geometry(uvw, freq, nx, ny, px, px, eps, True, "mixed", nbatch=nb, plan_opts={"sigma": 2.0})
    st = geom.stats()
    assert st["dense_enabled"] and 0.5 * st["nvis"] < st["dense_vis"] < st["nvis"]
    d = geom.ms2dirty(torch.tensor(ms, dtype=torch.complex64, device=dev)).double().cpu().numpy()
    v = geom.dirty2ms(torch.tensor(img, dtype=torch.float32, device=dev)).cpu().numpy()
    geom.close()
    for b in range(nb):
        ref_d = ox.explicit_ms2dirty(uvw, freq, ms[..., b], None, nx, ny, px, px, True)
        ref_v = ox.explicit_dirty2ms(uvw, freq, img[b], None, px, px, True)
        assert rel_l2(d[b], ref_d) < eps
        assert rel_l2(v[..., b].astype(numpy.complex128), ref_v) < eps


def test_unaligned_values_match_aligned(small_obs):
    """Visibility arrays that are not 16-byte aligned take the element-wise prologue and
    epilogue; results equal those of aligned arrays."""
    import torch

    uvw, _ = small_obs
    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(41)
    nrow, nb = uvw.shape[0], 2
    vals = rng.normal(size=(nrow, 2, nb)) + 1j * rng.normal(size=(nrow, 2, nb))
    flat = torch.zeros(nrow * 2 * nb + 1, dtype=torch.complex64, device=dev)
    flat[1:] = torch.tensor(vals.reshape(-1), dtype=torch.complex64, device=dev)
    ms_off = flat[1:].view(nrow, 2, nb)  # 8 bytes past an aligned start
    assert ms_off.data_ptr() % 16 == 8
    ms_al = ms_off.clone()
    assert ms_al.data_ptr() % 16 == 0
    img = torch.tensor(rng.normal(size=(nb, 64, 64)), dtype=torch.float32, device=dev)
    geom = wgridder.Geometry(uvw, FREQ2, 64, 64, 6e-4, 6e-4, 1e-5, True, "mixed", nbatch=nb)
    d_al = geom.ms2dirty(ms_al).clone()
    d_off = geom.ms2dirty(ms_off).clone()
    # the gridders flush with float atomics, so invert results of two calls differ at the
    # level of FP32 rounding in a different summation order; predict has no atomics
    assert (torch.linalg.norm(d_al - d_off) / torch.linalg.norm(d_al)).item() < 1e-6
    v_al = geom.dirty2ms(img).clone()
    out = torch.zeros_like(flat)
    geom.dirty2ms(img, out=out[1:].view(nrow, 2, nb))
    torch.cuda.synchronize()
    assert torch.equal(v_al.reshape(-1), out[1:])
    geom.close()


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_tile32_matches_explicit(prec):
    """32 x 32 sort tiles (two origin-digit passes, the T = 32 gatherer and degridder) on the
    dense-core geometry, against the brute-force sums."""
    import torch

    opts = {"tile": 32, "sigma": 2.0}
    uvw, freq = _dense_core(seed=35)
    rng = numpy.random.default_rng(36)
    nx, ny, px, eps = 64, 48, 2e-3, 1e-5
    ms = rng.normal(size=(uvw.shape[0], freq.size)) + 1j * rng.normal(size=(uvw.shape[0], freq.size))
    img = rng.normal(size=(nx, ny))
    plan = {}
    d = wgridder.ms2dirty(uvw, freq, ms, None, nx, ny, px, px, 0, 0, eps, True, precision=prec, plan_out=plan,
                          plan_opts=opts)
    assert plan["tile"] == 32 and plan["dense_items"] > 0
    v = wgridder.dirty2ms(uvw, freq, img, None, px, px, 0, 0, eps, True, precision=prec, plan_opts=opts)
    ref_d = ox.explicit_ms2dirty(uvw, freq, ms, None, nx, ny, px, px, True)
    ref_v = ox.explicit_dirty2ms(uvw, freq, img, None, px, px, True)
    assert rel_l2(numpy.asarray(d, numpy.float64), ref_d) < eps
    assert rel_l2(numpy.asarray(v, numpy.complex128), ref_v) < eps
