"""CPU-only tests: C-ABI library loads and exports the header's symbols, host-side plan
logic, data-model accessors, pol conversion vs the oracle matrices, dispatch errors."""

import ctypes
import math
import os
import re

import numpy
import pytest

import rascil_amd
from rascil_amd import _lib, wgridder
from rascil_amd.data_models import (BlockVisibility, Direction, PolarisationFrame, create_image,
                                    direction_to_lmn, get_parameter, image_is_canonical,
                                    lmn_to_direction)
from rascil_amd.imaging_base import normalise_sumwt, shift_vis_to_image
from rascil_amd.polarisation import convert_pol_frame
from oracle import explicit as ox

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    header = open(os.path.join(ROOT, "include", "hipng.h")).read()
    declared = set(re.findall(r"\b(hipng_[a-z0-9_]+)\s*\(", header))
    assert declared == set(_lib.EXPORTS)
    lib = ctypes.CDLL(_lib.LIBPATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert _lib.load().hipng_version() >= 100


def test_library_and_torch_share_one_hip_runtime():
    # load() in a fresh process, before anything imports torch: torch's HIP / HSA runtimes must
    # be the only ones mapped (a second copy makes device initialisation fail)
    import subprocess
    import sys
    pytest.importorskip("torch")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from rascil_amd import _lib\n_lib.load()\nimport torch\n"
            "print('\\n'.join(sorted(set(l.split()[-1] for l in open('/proc/self/maps')"
            " if 'libamdhip64' in l or 'libhsa-runtime64' in l))))\n"
            % ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr
    mapped = out.stdout.split()
    assert sum("libamdhip64" in m for m in mapped) == 1, mapped
    assert sum("libhsa-runtime64" in m for m in mapped) == 1, mapped

@pytest.mark.parametrize("eps,prec", [(1e-4, "mixed"), (1e-5, "mixed"), (1e-7, "fp64"),
                                      (1e-12, "fp64"), (1e-3, "fp32")])
@pytest.mark.parametrize("wst", [False, True])
def test_plan_params(eps, prec, wst):
    p = wgridder.plan_params(600.0, 550.0, 1.0, 400.0, 512, 512, 5e-4, 5e-4, eps, wst, prec)
    assert p.gx % p.tile == 0 and p.gy % p.tile == 0
    assert p.gx >= 1.2 * 512 and p.ntx == p.gx // p.tile and p.ntx >= 2
    assert 4 <= p.W <= 16 and p.tile >= p.W
    assert p.est_error <= eps
    if wst:
        # plane spacing keeps |dw ((n-1) - nshift)| <= 1/(2 sigma) over the image
        r2 = 2 * (256 * 5e-4) ** 2
        nm1 = -r2 / (numpy.sqrt(1 - r2) + 1)
        assert p.nshift == pytest.approx(nm1 / 2)
        assert p.dw * abs(nm1 / 2) <= 0.5 / p.sigma * (1 + 1e-12)
        # every |w| in [wmin, wmax] gets W planes inside [0, nplanes)
        for w in (1.0, 400.0):
            wh = (w - p.w0) / p.dw
            s = numpy.floor(wh - p.W / 2) + 1
            assert 0 <= s and s + p.W <= p.nplanes
    else:
        assert p.nplanes == 1
    assert p.nplanes * p.ntx * p.nty < 2 ** p.keybits + 1


def test_plan_errors():
    with pytest.raises(ValueError):
        wgridder.plan_params(1, 1, 0, 1, 511, 512, 1e-3, 1e-3, 1e-5, True)  # odd size
    with pytest.raises(ValueError):
        wgridder.plan_params(1, 1, 0, 1, 512, 512, 1e-3, 1e-3, 1e-9, True, "mixed")  # eps too small
    with pytest.raises(ValueError):
        wgridder.plan_params(1, 1, 0, 1, 2048, 2048, 1e-3, 1e-3, 1e-5, True)  # beyond the sphere


C2_EXTENT = (34955.4, 38229.1, 0.0, 15237.9)  # 512 stations, 8 channels 100-156 MHz


@pytest.mark.parametrize("prec,eps", [("mixed", 1e-6), ("mixed", 5e-6), ("fp32", 1e-4), ("fp32", 1e-3)])
def test_accuracy_contract_refuses_unreachable_epsilon(prec, eps):
    """Like ducc0 for single precision, an epsilon the mode cannot deliver raises instead of
    being silently missed: mixed below 1e-5 (its FP32 grid/FFT floor is measured at
    1.7e-6 - 3.2e-6 at 8192^2), FP32 below the phase error of FP32 grid coordinates at the
    image edge (3.9e-3 for the 8192^2, 2.2e-5 rad image of the 134M-visibility study)."""
    with pytest.raises(ValueError, match="epsilon"):
        wgridder.plan_params(*C2_EXTENT, 8192, 8192, 2.2e-5, 2.2e-5, eps, True, prec, 1, 134_000_000)


def test_accuracy_contract_floor_in_est_error():
    """Accepted plans report the modelled total error, and it stays below epsilon."""
    for prec, eps in [("fp64", 1e-6), ("mixed", 1e-5), ("fp32", 1e-2)]:
        p = wgridder.plan_params(*C2_EXTENT, 8192, 8192, 2.2e-5, 2.2e-5, eps, True, prec, 1, 134_000_000)
        assert p.est_error <= eps
    pm = wgridder.plan_params(*C2_EXTENT, 8192, 8192, 2.2e-5, 2.2e-5, 1e-5, True, "mixed", 1, 134_000_000)
    pd = wgridder.plan_params(*C2_EXTENT, 8192, 8192, 2.2e-5, 2.2e-5, 1e-5, True, "fp64", 1, 134_000_000)
    assert pm.est_error > pd.est_error  # the FP32 roundoff floor is part of the model
    # the FP32 coordinate floor grows with the image: a 64^2 image accepts 1e-4
    p = wgridder.plan_params(600.0, 550.0, 1.0, 400.0, 64, 64, 2e-3, 2e-3, 1e-4, True, "fp32")
    assert p.est_error <= 1e-4


def test_plan_options_pin_the_plan():
    base = dict(umax=600.0, vmax=550.0, wmin=1.0, wmax=400.0, npix_x=512, npix_y=512, pixsize_x=5e-4,
                pixsize_y=5e-4, epsilon=1e-5, do_wstacking=True, precision="mixed")
    auto = wgridder.plan_params(**base)
    for sg in (1.5, 2.0, 2.5):
        p = wgridder.plan_params(**base, plan_opts={"sigma": sg})
        assert p.sigma == sg and p.est_error <= 1e-5
    # at oversampling 1.25 the grid correction amplifies the FP32 transforms' roundoff ~1000x
    # at the image corner: that sigma cannot deliver 1e-5 in mixed precision, FP64 can
    with pytest.raises(ValueError, match="epsilon"):
        wgridder.plan_params(**base, plan_opts={"sigma": 1.25})
    assert wgridder.plan_params(**dict(base, precision="fp64"), plan_opts={"sigma": 1.25}).sigma == 1.25
    p = wgridder.plan_params(**base, plan_opts={"sigma": 2.0, "W": 9})
    assert p.W == 9
    p = wgridder.plan_params(**base, plan_opts={"tile": 32})
    assert p.tile == 32
    assert wgridder.plan_params(**base, plan_opts={"dense_min": -1}).dense_min == 0
    assert wgridder.plan_params(**base, plan_opts={"sigma": 2.0, "dense_min": 96}).dense_min == 96
    assert auto.dense_min in (0, 256)
    # an extent beyond the data's widens the plane range, never narrows it
    p = wgridder.plan_params(**base, plan_opts={"sigma": auto.sigma, "extent": (0.0, 0.0, 0.0, 800.0)})
    assert p.nplanes > auto.nplanes and p.umax == 600.0 and p.wmin == 0.0 and p.wmax == 800.0
    with pytest.raises(ValueError):
        wgridder.plan_params(**base, plan_opts={"sigma": 1.6})  # not a tabulated oversampling
    with pytest.raises(ValueError):
        wgridder.plan_params(**base, plan_opts={"sigma": 1.25, "W": 4})  # W too small for 1e-5
    with pytest.raises(ValueError):
        wgridder.plan_params(**base, plan_opts={"tile": 64})
    with pytest.raises(ValueError):
        wgridder.plan_params(**base, plan_opts={"oversampling": 2.0})  # unknown key


def test_plan_choice_is_kept_per_argument_set():
    # the library keeps its last plan choices: the same arguments give the same plan, evicted
    # ones are recomputed identically, options and epsilon are part of the key, and rejected
    # arguments are rejected on every call
    base = dict(umax=600.0, vmax=550.0, wmin=1.0, wmax=400.0, npix_x=512, npix_y=512, pixsize_x=5e-4,
                pixsize_y=5e-4, epsilon=1e-5, do_wstacking=True, precision="mixed", nvis=10_000)
    names = [f for f, _ in _lib.HipngPlan._fields_]

    def same(a, b):
        return all(getattr(a, f) == getattr(b, f) for f in names)

    first = wgridder.plan_params(**base)
    assert same(first, wgridder.plan_params(**base))
    for k in range(20):
        assert wgridder.plan_params(**dict(base, umax=600.0 + k)).umax == 600.0 + k
    assert same(first, wgridder.plan_params(**base))
    assert first.tile == 16 and wgridder.plan_params(**base, plan_opts={"tile": 32}).tile == 32
    assert wgridder.plan_params(**dict(base, epsilon=1e-3)).W < first.W
    for _ in range(2):
        with pytest.raises(ValueError, match="epsilon"):
            wgridder.plan_params(**base, plan_opts={"sigma": 1.25})


def test_pol_conversion_matches_oracle():
    rng = numpy.random.default_rng(0)
    v = rng.normal(size=(7, 3, 4)) + 1j * rng.normal(size=(7, 3, 4))
    for a, b in [("linear", "stokesIQUV"), ("stokesIQUV", "linear"), ("circular", "stokesIQUV"),
                 ("stokesIQUV", "circular"), ("linear", "stokesI"), ("stokesIQUV", "stokesI"),
                 ("linearnp", "stokesIQ"), ("stokesIQ", "linearnp"), ("circularnp", "stokesIV"),
                 ("stokesIV", "circularnp"), ("linearnp", "stokesI"), ("circularnp", "stokesI")]:
        nv = PolarisationFrame(a).npol
        v = rng.normal(size=(7, 3, nv)) + 1j * rng.normal(size=(7, 3, nv))
        got = convert_pol_frame(v, PolarisationFrame(a), PolarisationFrame(b), polaxis=2)
        exp = ox.convert_pol(v, a, b, polaxis=2)
        numpy.testing.assert_allclose(got, exp, atol=1e-15)
    with pytest.raises(ValueError):
        convert_pol_frame(v, PolarisationFrame("linear"), PolarisationFrame("circular"), polaxis=2)


def test_direction_roundtrip_and_lmn():
    pc = Direction(0.3, -0.7)
    d = lmn_to_direction(0.01, -0.02, pc)
    l, m, nm1 = direction_to_lmn(d, pc)
    assert (l, m) == pytest.approx((0.01, -0.02), abs=1e-14)
    assert nm1 == pytest.approx(numpy.sqrt(1 - 0.01 ** 2 - 0.02 ** 2) - 1, abs=1e-14)


def test_direction_to_lmn_closed_forms():
    """Spherical-trigonometry special cases (coordinate_support.py:446-465 semantics): an RA
    offset delta at declination d gives l = cos d sin delta > 0 (l increases east-wards, with
    RA), m = sin d cos d (1 - cos delta); a declination offset gives l = 0, m = sin delta;
    on the equator n - 1 = cos delta - 1; at the north pole l = cos dec sin ra,
    m = -cos dec cos ra."""
    d0, delta = math.radians(-35.0), math.radians(1.0)
    l, m, nm1 = direction_to_lmn(Direction(math.pi + delta, d0), Direction(math.pi, d0))
    assert l > 0
    assert l == pytest.approx(math.cos(d0) * math.sin(delta), abs=1e-15)
    assert m == pytest.approx(math.sin(d0) * math.cos(d0) * (1 - math.cos(delta)), abs=1e-15)
    assert nm1 == pytest.approx(math.sin(d0) ** 2 + math.cos(d0) ** 2 * math.cos(delta) - 1, abs=1e-15)
    l, m, nm1 = direction_to_lmn(Direction(0.4, d0 + delta), Direction(0.4, d0))
    assert (l, m) == pytest.approx((0.0, math.sin(delta)), abs=1e-15)
    assert nm1 == pytest.approx(math.cos(delta) - 1, abs=1e-15)
    l, m, nm1 = direction_to_lmn(Direction(delta, 0.0), Direction(0.0, 0.0))
    assert (l, m, nm1) == pytest.approx((math.sin(delta), 0.0, math.cos(delta) - 1), abs=1e-15)
    ra, dec = 0.7, math.radians(80.0)
    l, m, _ = direction_to_lmn(Direction(ra, dec), Direction(0.0, math.pi / 2))
    assert (l, m) == pytest.approx((math.cos(dec) * math.sin(ra), -math.cos(dec) * math.cos(ra)), abs=1e-15)


@pytest.mark.parametrize("comp_offset", [(1.0, 0.0), (1.3, 0.4)])
def test_phase_rotation_tangent_false_matches_direct_prediction(comp_offset):
    """test_dft_skycomponent_visibility.py:51-132 restated on synthetic stations: the DFT of a
    component with phase centre (180, -35) deg, phase rotated with uvw re-projection to
    (181, -35) deg, equals the DFT computed directly with phase centre (181, -35) at hour
    angles shifted by -1 deg, vis and uvw (rtol 3e-6 in the reference). With the component at
    the new centre (the reference's case) the rotated vis are constant; a component elsewhere
    also ties the phasor's sign to the l axis of the uvw synthesis (east = +u)."""
    from rascil_amd import simulation
    from rascil_amd.imaging_base import phaserotate_visibility

    rng = numpy.random.default_rng(3)
    ants = simulation.gaussian_disc_stations(25, 100.0, 300.0, rng)
    ha = (numpy.pi / 43200.0) * numpy.arange(0.0, 300.0, 100.0)
    freq = numpy.linspace(1.0e8, 1.1e8, 3)
    pc = Direction(math.pi, math.radians(-35.0))
    newpc = Direction(math.radians(181.0), math.radians(-35.0))
    comp = Direction(math.radians(180.0 + comp_offset[0]), math.radians(-35.0 + comp_offset[1]))
    flux = numpy.array([[100.0], [80.0], [60.0]])

    def dft(bvis):
        l, m, nm1 = direction_to_lmn(comp, bvis.phasecentre)
        return ox.dft_cpu_looped(numpy.array([[l, m, nm1]]), bvis.uvw_lambda, flux[None])

    vis = simulation.create_blockvisibility(ants, ha, freq, pc)
    vis.vis = dft(vis)
    direct = simulation.create_blockvisibility(ants, ha - (newpc.ra - pc.ra), freq, newpc)
    direct.vis = dft(direct)
    rot = phaserotate_visibility(vis, newpc, tangent=False)
    assert rot.phasecentre == newpc
    numpy.testing.assert_allclose(rot.uvw, direct.uvw, rtol=3e-6, atol=1e-9)
    numpy.testing.assert_allclose(rot.vis, direct.vis, rtol=3e-6, atol=1e-9)
    if comp_offset == (1.0, 0.0):
        numpy.testing.assert_allclose(rot.vis[..., 0], numpy.broadcast_to(flux[:, 0], rot.vis.shape[:3]),
                                      rtol=1e-9)
    # the tangent-plane rotation keeps uvw and phase centre, as on the imaging path
    tan = phaserotate_visibility(vis, newpc, tangent=True)
    assert numpy.array_equal(tan.uvw, vis.uvw) and tan.phasecentre == pc


def test_blockvisibility_accessors_and_image():
    pf = PolarisationFrame("linear")
    vis = numpy.ones((2, 3, 4, 4), "c16")
    flags = numpy.zeros(vis.shape, int)
    flags[0, 1] = 1
    b = BlockVisibility(vis, numpy.ones((2, 3, 3)), [1e8, 1.1e8, 1.2e8, 1.3e8], pf,
                        Direction(0, -0.5), flags=flags)
    assert b.blockvisibility_acc.flagged_vis[0, 1].sum() == 0
    assert b.blockvisibility_acc.flagged_imaging_weight[0, 0].sum() == 16
    assert b["vis"].data is b.vis
    uvl = b.uvw_lambda
    assert uvl.shape == (2, 3, 4, 3)
    assert uvl[0, 0, 1, 0] == pytest.approx(1.1e8 / 299792458.0)
    im = create_image(64, 1e-3, Direction(0, -0.5), [1e8, 1.1e8], PolarisationFrame("stokesIQUV"))
    assert image_is_canonical(im) and im.pixels.shape == (2, 4, 64, 64)
    assert list(numpy.round(im.image_acc.wcs.freq_to_pix([1e8, 1.1e8])).astype(int)) == [0, 1]
    assert im.wcs.crpix[0] == 33 and im.wcs.cdelt[0] < 0


def test_normalise_and_shift_noop():
    im = create_image(16, 1e-3, Direction(0, -0.5), [1e8], PolarisationFrame("stokesIQ"))
    im.pixels[...] = 2.0
    out = normalise_sumwt(im, numpy.array([[4.0, 0.0]]))
    assert numpy.all(out.pixels[0, 0] == 0.5) and numpy.all(out.pixels[0, 1] == 0.0)
    b = BlockVisibility(numpy.ones((1, 2, 1, 1), "c16"), numpy.ones((1, 2, 3)), [1e8],
                        PolarisationFrame("stokesI"), Direction(0, -0.5))
    assert shift_vis_to_image(b, im) is b


def test_shift_vis_to_image_rotates_phase():
    pc_vis = Direction(0.0, -0.5)
    im = create_image(16, 1e-3, Direction(0.01, -0.49), [1e8], PolarisationFrame("stokesI"))
    b = BlockVisibility(numpy.ones((1, 2, 1, 1), "c16"), numpy.array([[[10., 20., 5.], [-3., 4., 1.]]]),
                        [1e8], PolarisationFrame("stokesI"), pc_vis)
    s = shift_vis_to_image(b, im, inverse=False)
    l, m, nm1 = direction_to_lmn(im.phasecentre, pc_vis)
    ph = ox.phasor(b.uvw_lambda, l, m, nm1)
    numpy.testing.assert_allclose(s.vis[..., 0], numpy.conj(ph), atol=1e-14)
    assert s.phasecentre == im.phasecentre


def test_dispatch_unknown_context():
    from rascil_amd.imaging import invert_blockvisibility, predict_blockvisibility

    with pytest.raises(ValueError):
        predict_blockvisibility(None, None, context="nope")
    with pytest.raises(ValueError):
        invert_blockvisibility(None, None, context="nope")


def test_get_parameter():
    assert get_parameter({"epsilon": 1e-5}, "epsilon", 1e-12) == 1e-5
    assert get_parameter({}, "epsilon", 1e-12) == 1e-12
    assert get_parameter(None, "x", 3) == 3
    # a key that is present is returned as given, None included (parameters.py:119-125)
    assert get_parameter({"epsilon": None}, "epsilon", 1e-12) is None


def test_image_with_moment_axis_is_canonical():
    """image/operations.py:96-98 accepts FREQ or MOMENT as the fourth axis."""
    im = create_image(32, 1e-3, Direction(0.0, -0.5), [1e8])
    assert image_is_canonical(im)
    im.wcs.ctype[3] = "MOMENT"
    assert image_is_canonical(im)
    im.wcs.ctype[3] = "VELO"
    assert not image_is_canonical(im)


def test_product_has_no_cpu_fallback(monkeypatch):
    """Without the library the product raises instead of computing anything."""
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIBPATH", "/nonexistent/_hipng.so")
    with pytest.raises(_lib.HipngLibraryMissing):
        _lib.load()


def test_package_does_not_import_oracle():
    for dirpath, _, files in os.walk(os.path.join(ROOT, "rascil_amd")):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", src, re.M), f


def test_copy_visibility_is_deep_and_zeroes_on_request():
    """visibility/base.py:105-118: a deep copy; zero=True clears vis only."""
    import math

    from rascil_amd import Direction, PolarisationFrame, copy_visibility, simulation

    rng = numpy.random.default_rng(2)
    ants = simulation.gaussian_disc_stations(6, 100.0, 300.0, rng)
    bvis = simulation.create_blockvisibility(ants, numpy.array([0.0, 0.1]), [1e8, 1.1e8],
                                             Direction(0.0, math.radians(-30)), PolarisationFrame("linear"))
    bvis.vis = rng.normal(size=bvis.vis.shape) + 1j * rng.normal(size=bvis.vis.shape)
    c = copy_visibility(bvis)
    numpy.testing.assert_array_equal(c.vis, bvis.vis)
    c.vis[0, 0, 0, 0] = 99.0
    c.weight[...] = 7.0
    assert bvis.vis[0, 0, 0, 0] != 99.0 and not numpy.any(bvis.weight == 7.0)
    z = copy_visibility(bvis, zero=True)
    assert not z.vis.any()
    numpy.testing.assert_array_equal(z.weight, bvis.weight)
    numpy.testing.assert_array_equal(z.uvw, bvis.uvw)
    assert bvis.vis.any()
