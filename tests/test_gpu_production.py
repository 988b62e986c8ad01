"""GPU parity at the production shapes, through the C-ABI:

* the plan the 4.29e9-visibility benchmark runs (W 8, sigma 1.5, 12288^2 grid, 57 planes,
  dense-run path on) pinned through the plan options on a time-contiguous subsample of its
  baselines, against FP64 direct sums (the HIP FP64 DFT, itself checked against the oracle
  in test_gpu_ng.py, and the FP64 adjoint sum);
* the out-of-place FFT plans at those lengths, checked for leaving their input untouched;
* the precision study geometry (1.34e8 visibilities): every (mode, epsilon) the accuracy
  contract accepts meets epsilon, and the ones it cannot meet raise;
* 60 uneven snapshot jobs through prebuilt geometries and graph replay vs the oracle's
  brute-force sums;
* gradients of a gain-calibration chi-squared at 2048^2 with 1M rows x 4 channels and 256
  antenna gains: finite differences and the FP64 adjoint.
"""

import math

import numpy
import pytest

from oracle import explicit as ox
from rascil_amd import simulation, wgridder

pytestmark = pytest.mark.gpu

C = 299792458.0
SCALES = numpy.array([1.0, 0.5, -0.25, 0.125])


def _extent(uvw, freq):
    a = numpy.abs(uvw).max(axis=0)
    return (float(a[0] * freq.max() / C), float(a[1] * freq.max() / C),
            float(numpy.abs(uvw[:, 2]).min() * freq.min() / C), float(a[2] * freq.max() / C))


def _sources(N, cell, nsrc, rng, radius=None):
    """Pixel indices of nsrc sources (uniform in the inner 3/4, or within `radius` rad)."""
    if radius is None:
        ix = rng.integers(N // 8, N - N // 8, nsrc)
        iy = rng.integers(N // 8, N - N // 8, nsrc)
    else:
        r = radius * numpy.sqrt(rng.uniform(0, 1, nsrc))
        phi = rng.uniform(0, 2 * numpy.pi, nsrc)
        ix = numpy.clip(N // 2 + numpy.round(r * numpy.cos(phi) / cell).astype(int), 0, N - 1)
        iy = numpy.clip(N // 2 + numpy.round(r * numpy.sin(phi) / cell).astype(int), 0, N - 1)
    return ix, iy


def _distinct(ix, iy, flux):
    """Merge sources that share a pixel (their fluxes add)."""
    pix, inv = numpy.unique(numpy.stack([ix, iy], 1), axis=0, return_inverse=True)
    f = numpy.zeros(len(pix))
    numpy.add.at(f, inv.ravel(), flux)
    return pix[:, 0], pix[:, 1], f


class DirectSums:
    """FP64 references on the GPU for pixel-centred sources (x = (ix - N/2) cell, ducc0 sign
    convention with uvw as passed to ms2dirty / dirty2ms): vis(row, chan) = sum_k F_k / n_k
    exp(-2 pi i f/c (u x_k + v y_k - w (n_k - 1))) by the HIP FP64 DFT, and the adjoint
    dirty(x_k) = Re sum_i w_i V_i exp(+2 pi i f/c (...)) / n_k by torch in FP64."""

    def __init__(self, uvw_d, freq, ix, iy, N, cell, chunk_rows=1 << 19):
        import torch

        self.torch = torch
        self.dev = uvw_d.device
        self.uvw_d, self.freq = uvw_d, numpy.asarray(freq, "f8")
        x, y = (ix - N // 2) * cell, (iy - N // 2) * cell
        self.n = numpy.sqrt(1 - x * x - y * y)
        self.dc = torch.tensor(numpy.stack([-x, y, self.n - 1], axis=1), dtype=torch.float64, device=self.dev)
        self.sgn = torch.tensor([-1.0, 1.0, -1.0], dtype=torch.float64, device=self.dev)
        self.fc = torch.tensor(self.freq / C, dtype=torch.float64, device=self.dev)
        self.chunk = chunk_rows

    def _uvl(self, r0, r1):
        return ((self.uvw_d[r0:r1] * self.sgn)[:, None, :] * self.fc[None, :, None]).contiguous()

    def predict(self, flux):
        """[nrow, nchan] complex128 of the sources with fluxes `flux` (the pixel values)."""
        from rascil_amd.dft import dft_device

        torch = self.torch
        nrow, nchan = self.uvw_d.shape[0], len(self.freq)
        fl = torch.tensor(flux / self.n, dtype=torch.complex128, device=self.dev)
        fl = fl.reshape(-1, 1, 1).expand(-1, nchan, 1).contiguous()
        out = torch.empty((nrow, nchan), dtype=torch.complex128, device=self.dev)
        for r0 in range(0, nrow, self.chunk):
            r1 = min(nrow, r0 + self.chunk)
            out[r0:r1] = dft_device(self.dc, fl, self._uvl(r0, r1).reshape(1, r1 - r0, nchan, 3))[0, :, :, 0]
        return out

    def adjoint(self, vis, wgt=None):
        """[nb, K] FP64 adjoint sums at the source pixels of vis [nrow, nchan, nb]."""
        torch = self.torch
        nrow = self.uvw_d.shape[0]
        dcs = self.dc.cpu().numpy()
        out = torch.zeros((vis.shape[2], dcs.shape[0]), dtype=torch.float64, device=self.dev)
        for r0 in range(0, nrow, self.chunk):
            r1 = min(nrow, r0 + self.chunk)
            uvl = self._uvl(r0, r1)
            V = vis[r0:r1].to(torch.complex128)
            if wgt is not None:
                V = V * wgt[r0:r1].double()
            for k in range(dcs.shape[0]):
                ph = uvl[..., 0] * dcs[k, 0] + uvl[..., 1] * dcs[k, 1] + uvl[..., 2] * dcs[k, 2]
                out[:, k] += (V * torch.polar(torch.ones_like(ph), 2 * math.pi * ph)[..., None]).real.sum(dim=(0, 1))
        return out / torch.tensor(self.n, dtype=torch.float64, device=self.dev)


def _max_rel(got, ref):
    return float((got - ref).abs().max() / ref.abs().max())


# --------------------------------------------------------------------------------------
# the benchmark's plan on a subsample of its baselines
# --------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c1_subsample():
    uvw, freq, _ = simulation.config1_geometry()
    full_extent = _extent(uvw, freq)
    ntimes, nbl = 128, uvw.shape[0] // 128
    rng = numpy.random.default_rng(21)
    lengths = numpy.linalg.norm(uvw[:nbl, :2], axis=1)
    bl = numpy.unique(numpy.concatenate([rng.choice(nbl, 1000, replace=False), numpy.argsort(lengths)[-8:],
                                         numpy.argsort(lengths)[:48]]))
    rows = (numpy.arange(ntimes)[:, None] * nbl + bl[None, :]).reshape(-1)  # all times of each baseline
    return numpy.ascontiguousarray(uvw[rows]), freq, full_extent


def test_benchmark_plan_meets_epsilon_on_subsample(c1_subsample):
    """The production plan (4 polarisations, mixed precision, epsilon 1e-5) pinned with the
    plan options (sigma 1.5 and the full geometry's uv extent) on 1,024 baselines x 128 times x
    64 channels: predict and invert of 20 pixel-centred sources meet 1e-5 against FP64 direct
    sums, with a large share of the visibilities on the dense-run path."""
    import torch

    uvw, freq, ext = c1_subsample
    dev = torch.device("cuda", 0)
    N, cell, eps, nb = 8192, 1.1e-5, 1e-5, 4
    opts = {"sigma": 1.5, "extent": ext}
    rng = numpy.random.default_rng(22)
    ix, iy = _sources(N, cell, 20, rng)
    ix, iy, flux = _distinct(ix, iy, rng.uniform(0.1, 1.0, 20))
    uvw_d = torch.tensor(uvw, device=dev)
    ds = DirectSums(uvw_d, freq, ix, iy, N, cell)
    model = torch.zeros((nb, N, N), dtype=torch.float32, device=dev)
    for b in range(nb):
        model[b, ix, iy] = torch.tensor(flux * SCALES[b], dtype=torch.float32, device=dev)
    plan = {}
    vis = wgridder.dirty2ms(uvw_d, freq, model, None, cell, cell, 0, 0, eps, True, precision="mixed",
                            plan_out=plan, plan_opts=opts)
    assert (plan["W"], plan["sigma"], plan["gx"], plan["gy"], plan["nplanes"], plan["tile"], plan["dense_min"]) == \
        (8, 1.5, 12288, 12288, 57, 16, 256)
    assert plan["dense_items"] > 0.15 * uvw.shape[0] * len(freq)  # both gridding paths carry work
    ref = ds.predict(flux)
    sc = torch.tensor(SCALES, dtype=torch.complex128, device=dev)
    err = _max_rel(vis.to(torch.complex128), ref[..., None] * sc)
    print(f"benchmark plan predict max rel {err:.2e}")
    assert err <= eps
    ms = (ref[..., None] * sc).to(torch.complex64)
    wgt = torch.empty(ms.shape, dtype=torch.float32, device=dev).uniform_(0.5, 1.5)
    del vis, ref
    dirty = wgridder.ms2dirty(uvw_d, freq, ms, wgt, N, N, cell, cell, 0, 0, eps, True, precision="mixed",
                              plan_opts=opts)
    got = dirty[:, ix, iy].double()
    exact = ds.adjoint(ms, wgt)
    err = _max_rel(got, exact)
    print(f"benchmark plan invert max rel {err:.2e}")
    assert err <= eps


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fft_plans_keep_input_at_production_lengths(c1_subsample, prec):
    """At the production transform lengths (rows of 12288, the band of the full uv extent)
    the geometry checks each out-of-place rocFFT plan on a pattern; whichever way that goes
    (input kept: buffers cleared once per call; input overwritten: cleared per pass) invert and
    predict of a few sources meet epsilon."""
    import torch

    uvw, freq, ext = c1_subsample
    dev = torch.device("cuda", 0)
    N, cell, eps = 8192, 1.1e-5, 1e-5
    sel = numpy.arange(0, uvw.shape[0], 97)
    u = torch.tensor(uvw[sel], device=dev)
    geom = wgridder.Geometry(u, freq[::8], N, N, cell, cell, eps, True, prec, nbatch=1,
                             plan_opts={"sigma": 1.5, "extent": ext})
    st = geom.stats()
    print(f"{prec}: out-of-place FFT plans keep their input: {st['fft_keeps_input']}")
    assert geom.plan["gx"] == 12288
    rng = numpy.random.default_rng(23)
    ix, iy = _sources(N, cell, 6, rng)
    ix, iy, flux = _distinct(ix, iy, rng.uniform(0.1, 1.0, 6))
    ds = DirectSums(u, freq[::8], ix, iy, N, cell)
    rd = torch.float64 if prec == "fp64" else torch.float32
    cd = torch.complex128 if prec == "fp64" else torch.complex64
    img = torch.zeros((1, N, N), dtype=rd, device=dev)
    img[0, ix, iy] = torch.tensor(flux, dtype=rd, device=dev)
    ref = ds.predict(flux)
    for _ in range(2):  # a second call sees the buffers the first one left
        vis = geom.dirty2ms(img)
        assert _max_rel(vis[..., 0].to(torch.complex128), ref) <= eps
        dirty = geom.ms2dirty(ref.to(cd).reshape(-1, ref.shape[1], 1).contiguous())
        exact = ds.adjoint(ref.to(cd).reshape(-1, ref.shape[1], 1))
        assert _max_rel(dirty[:, ix, iy].double(), exact) <= eps
    geom.close()


# --------------------------------------------------------------------------------------
# precision study geometry: 16.7M rows x 8 channels, 200 sources within 5 deg
# --------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2():
    import torch

    uvw, freq, _ = simulation.config1_geometry(nchan=8)
    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(3)
    N, cell, nsrc = 8192, 2.2e-5, 200
    ix, iy = _sources(N, cell, nsrc, rng, radius=math.radians(5.0))
    flux = simulation.powerlaw_fluxes(nsrc, -1.6, 1e-3, 1.0, rng)
    uvw_d = torch.tensor(uvw, device=dev)
    model = numpy.zeros((N, N))
    numpy.add.at(model, (ix, iy), flux)
    # distinct pixels (a few sources share one); the model's value there is their sum
    pix = numpy.unique(numpy.stack([ix, iy], 1), axis=0)
    ds = DirectSums(uvw_d, freq, pix[:, 0], pix[:, 1], N, cell, chunk_rows=1 << 20)
    vtrue = ds.predict(model[pix[:, 0], pix[:, 1]])
    probe = numpy.arange(0, len(pix), 10)
    dsp = DirectSums(uvw_d, freq, pix[probe, 0], pix[probe, 1], N, cell, chunk_rows=1 << 20)
    exact = dsp.adjoint(vtrue[..., None])[0]
    return dict(uvw_d=uvw_d, freq=freq, N=N, cell=cell, model=model, vtrue=vtrue, pix=pix[probe], exact=exact)


@pytest.mark.parametrize("prec,eps", [("fp64", 1e-3), ("fp64", 1e-4), ("fp64", 1e-5), ("mixed", 1e-3),
                                      ("mixed", 1e-4), ("mixed", 1e-5), ("fp32", 1e-2)])
def test_precision_study_accepted_modes_meet_epsilon(c2, prec, eps):
    """1.34e8 visibilities, 8192^2 at 2.2e-5 rad: predicted visibilities (all of them) and the
    dirty image at probe source pixels against FP64 direct sums, max relative error <= eps."""
    import torch

    dev = c2["uvw_d"].device
    rd = torch.float64 if prec == "fp64" else torch.float32
    cd = torch.complex128 if prec == "fp64" else torch.complex64
    img = torch.tensor(c2["model"], dtype=rd, device=dev)
    vis = wgridder.dirty2ms(c2["uvw_d"], c2["freq"], img, None, c2["cell"], c2["cell"], 0, 0, eps, True,
                            precision=prec)
    e_vis = _max_rel(vis.to(torch.complex128), c2["vtrue"])
    del vis
    dirty = wgridder.ms2dirty(c2["uvw_d"], c2["freq"], c2["vtrue"].to(cd), None, c2["N"], c2["N"], c2["cell"],
                              c2["cell"], 0, 0, eps, True, precision=prec)
    got = dirty[c2["pix"][:, 0], c2["pix"][:, 1]].double()
    e_img = _max_rel(got, c2["exact"])
    print(f"{prec} eps={eps:g}: vis max rel {e_vis:.2e}, dirty max rel {e_img:.2e}")
    assert e_vis <= eps and e_img <= eps


@pytest.mark.parametrize("prec,eps", [("mixed", 1e-6), ("fp32", 1e-3), ("fp32", 1e-4)])
def test_precision_study_unreachable_epsilon_raises(c2, prec, eps):
    """Mixed at 1e-6 (measured 3.2e-6 here) and FP32 below its coordinate floor (3.9e-3 at
    this image and uv extent) are refused instead of silently missed."""
    import torch

    rd = torch.float32
    img = torch.tensor(c2["model"], dtype=rd, device=c2["uvw_d"].device)
    with pytest.raises(ValueError, match="epsilon"):
        wgridder.dirty2ms(c2["uvw_d"], c2["freq"], img, None, c2["cell"], c2["cell"], 0, 0, eps, True,
                          precision=prec)


# --------------------------------------------------------------------------------------
# uneven snapshot jobs through prebuilt geometries and graph replay
# --------------------------------------------------------------------------------------
def test_uneven_jobs_graph_replay_match_explicit():
    """60 snapshot jobs with log-uniform row counts, image sizes 32-64, 1-3 channels, random
    declinations and 2-20 minute windows (mixed precision, epsilon 1e-4): each job's
    invert + predict is captured in a graph once and replayed; every replay equals the
    brute-force sums to epsilon."""
    import torch

    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(31)
    ants = simulation.core_plus_arms_stations(64, 0.7, 300.0, 500.0, 3000.0, rng)
    stream = torch.cuda.Stream()
    worst = 0.0
    for job in range(60):
        nrow_target = int(numpy.exp(rng.uniform(numpy.log(200), numpy.log(2000))))
        N = int(rng.choice([32, 48, 64]))
        nchan = int(rng.integers(1, 4))
        dec = math.radians(rng.uniform(-80, 10))
        window = rng.uniform(2, 20) / 60.0 * math.pi / 12
        a1, _ = simulation.baseline_indices(len(ants), autocorrelations=False)
        ntimes = max(1, nrow_target // len(a1) + 1)
        ha0 = rng.uniform(-2, 2) * math.pi / 12
        uvw = simulation.baseline_uvw(ants, ha0 + numpy.linspace(0, window, ntimes), dec,
                                      autocorrelations=False).reshape(-1, 3)[:nrow_target]
        freq = 1e8 + 2e6 * numpy.arange(nchan)
        umax = numpy.abs(uvw[:, :2]).max() * freq.max() / C
        cell = 0.4 / umax / 2  # Nyquist for the longest baseline, with margin
        eps = 1e-4
        ms = (rng.normal(size=(uvw.shape[0], nchan)) + 1j * rng.normal(size=(uvw.shape[0], nchan)))
        img = rng.normal(size=(N, N))
        geom = wgridder.Geometry(uvw, freq, N, N, cell, cell, eps, True, "mixed")
        ms_d = torch.tensor(ms, dtype=torch.complex64, device=dev)
        img_d = torch.tensor(img, dtype=torch.float32, device=dev)
        d_out = torch.empty((1, N, N), dtype=torch.float32, device=dev)
        v_out = torch.empty((uvw.shape[0], nchan, 1), dtype=torch.complex64, device=dev)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.stream(stream):
            geom.ms2dirty(ms_d, out=d_out)  # warm-up outside the capture
            geom.dirty2ms(img_d, out=v_out)
            stream.synchronize()
            with torch.cuda.graph(graph, stream=stream):
                geom.ms2dirty(ms_d, out=d_out)
                geom.dirty2ms(img_d, out=v_out)
        ref_d = ox.explicit_ms2dirty(uvw, freq, ms, None, N, N, cell, cell, True)
        ref_v = ox.explicit_dirty2ms(uvw, freq, img, None, cell, cell, True)
        for _ in range(2):
            d_out.zero_()
            v_out.zero_()
            graph.replay()
            torch.cuda.synchronize()
            ed = float(numpy.linalg.norm(d_out[0].double().cpu().numpy() - ref_d) / numpy.linalg.norm(ref_d))
            ev = float(numpy.linalg.norm(v_out[..., 0].cpu().numpy() - ref_v) / numpy.linalg.norm(ref_v))
            worst = max(worst, ed, ev)
            assert ed <= eps and ev <= eps, (job, uvw.shape[0], N, nchan, ed, ev)
        del graph
        geom.close()
    print(f"60 jobs, worst relative L2 error {worst:.2e}")


# --------------------------------------------------------------------------------------
# differentiable predict at 2048^2, 1M rows x 4 channels, 256 antenna gains
# --------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c4():
    import torch

    from rascil_amd import simulation as sim

    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(11)
    nant = 256
    ants = sim.core_plus_arms_stations(nant, 0.7, 500.0, 1000.0, 40000.0, rng)
    a1, a2 = sim.baseline_indices(nant, autocorrelations=False)
    ntimes = int(numpy.ceil(1_000_000 / len(a1)))
    uvw = sim.baseline_uvw(ants, numpy.linspace(-2, 2, ntimes) * numpy.pi / 12, math.radians(-30.0),
                           autocorrelations=False).reshape(-1, 3)
    freq = numpy.array([100e6, 110e6, 120e6, 130e6])
    N, cell, nsrc = 2048, 2e-5, 500
    ix, iy = _sources(N, cell, nsrc, rng)
    flux = sim.powerlaw_fluxes(nsrc, -1.6, 1e-3, 1.0, rng)
    model = numpy.zeros((N, N))
    numpy.add.at(model, (ix, iy), flux)
    uvw_d = torch.tensor(uvw, device=dev)
    pix = numpy.unique(numpy.stack([ix, iy], 1), axis=0)
    ds = DirectSums(uvw_d, freq, pix[:, 0], pix[:, 1], N, cell)
    vtrue = ds.predict(model[pix[:, 0], pix[:, 1]])
    gtrue = rng.uniform(0.9, 1.1, nant) * numpy.exp(1j * numpy.radians(rng.uniform(-30, 30, nant)))
    i1 = torch.tensor(numpy.tile(a1, ntimes), device=dev)
    i2 = torch.tensor(numpy.tile(a2, ntimes), device=dev)
    g_t = torch.tensor(gtrue, dtype=torch.complex128, device=dev)
    noise = torch.tensor(rng.normal(size=(2,) + tuple(vtrue.shape)), dtype=torch.float64, device=dev)
    vobs = (g_t[i1] * torch.conj(g_t[i2]))[:, None] * vtrue + 0.01 * torch.complex(noise[0], noise[1])
    return dict(uvw_d=uvw_d, freq=freq, N=N, cell=cell, model=model, ix=ix, iy=iy, vobs=vobs, i1=i1, i2=i2,
                gtrue=gtrue, nant=nant, rng=rng)


@pytest.mark.parametrize("fused", [False, True])
@pytest.mark.parametrize("prec,eps,tol,hg", [("fp64", 1e-10, 1e-6, 1e-5), ("mixed", 1e-5, 1e-3, 1e-2)])
def test_gain_calibration_gradients_at_production_size(c4, prec, eps, tol, hg, fused):
    """Loss = sum |V_obs - g_i conj(g_j) predict(I)|^2 over 1M rows x 4 channels: the image
    gradient (exact central differences: the loss is quadratic in the image) and the gain
    gradients (central differences) within 1e-6 (FP64) / 1e-3 (mixed), and the image gradient
    equal to the FP64 adjoint of the residual at source pixels; with the gains as torch
    operations around the degridder or fused into it (and into the gridder of backward)."""
    import torch

    from rascil_amd.autograd import apply_gains, predict

    dev = c4["uvw_d"].device
    rd = torch.float64 if prec == "fp64" else torch.float32
    cd = torch.complex128 if prec == "fp64" else torch.complex64
    N, cell = c4["N"], c4["cell"]
    geom = wgridder.Geometry(c4["uvw_d"], c4["freq"], N, N, cell, cell, eps, True, prec)
    vo, i1, i2 = c4["vobs"], c4["i1"], c4["i2"]

    j1, j2 = i1.to(torch.int32), i2.to(torch.int32)

    def loss(img, g):
        if fused:
            r = vo - predict(img, geometry=geom, gains=g, ant1=j1, ant2=j2).to(torch.complex128)
        else:
            v = predict(img, geometry=geom).to(torch.complex128)
            r = vo - apply_gains(v, g.to(torch.complex128), i1, i2)
        return (r.abs() ** 2).sum()

    rng = numpy.random.default_rng(12)
    img0 = torch.tensor(c4["model"] * 0.9, dtype=rd, device=dev)
    g0 = torch.tensor(c4["gtrue"] * (1 + 0.05 * rng.normal(size=c4["nant"])), dtype=cd, device=dev)
    img = img0.clone().requires_grad_(True)
    g = g0.clone().requires_grad_(True)
    loss(img, g).backward()
    gi, gg = img.grad.detach(), g.grad.detach()
    with torch.no_grad():
        for k in range(3):
            px, py = int(c4["ix"][k]), int(c4["iy"][k])
            ip, im_ = img0.clone(), img0.clone()
            ip[px, py] += 1.0
            im_[px, py] -= 1.0
            fd = (loss(ip, g0) - loss(im_, g0)).item() / 2.0
            assert abs(fd - gi[px, py].item()) <= tol * max(1.0, abs(fd)), (k, fd, gi[px, py].item())
        for k in (0, 7):
            for dirn in (1.0, 1j):
                gp, gm = g0.clone(), g0.clone()
                gp[k] += hg * dirn
                gm[k] -= hg * dirn
                fd = (loss(img0, gp) - loss(img0, gm)).item() / (2 * hg)
                an = gg[k].real.item() if dirn == 1.0 else gg[k].imag.item()
                assert abs(fd - an) <= tol * max(1.0, abs(fd)), (k, dirn, fd, an)
        # dL/dI(x_k) = Re sum_i conj(-2 conj(gprod_i) r_i)... = the FP64 adjoint of the residual
        v = predict(img0, geometry=geom).to(torch.complex128)
        gprod = g0.to(torch.complex128)[i1] * torch.conj(g0.to(torch.complex128))[i2]
        back = -2.0 * torch.conj(gprod)[:, None] * (vo - gprod[:, None] * v)
        pix = numpy.stack([c4["ix"][:8], c4["iy"][:8]], 1)
        ds = DirectSums(c4["uvw_d"], c4["freq"], pix[:, 0], pix[:, 1], N, cell)
        adj = ds.adjoint(back[..., None])[0]
        got = gi[pix[:, 0], pix[:, 1]].double()
        err = _max_rel(got, adj)
        print(f"{prec}: image gradient vs FP64 adjoint max rel {err:.2e}")
        assert err <= (1e-9 if prec == "fp64" else 1e-4)
    geom.close()


def test_tuned_concurrent_batch_matches_explicit():
    """40 uneven snapshot jobs through rascil_amd.batch: plan options tuned per job class on
    the GPU (every chosen plan meets epsilon), the jobs' invert + predict graphs replayed
    concurrently over 4 streams, twice; every job equals the brute-force sums to epsilon."""
    import torch

    from rascil_amd.batch import AutoTuner, Batch

    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(41)
    ants = simulation.core_plus_arms_stations(48, 0.7, 300.0, 500.0, 3000.0, rng)
    eps = 1e-4
    jobs, refs = [], []
    for job in range(40):
        nrow = int(numpy.exp(rng.uniform(numpy.log(200), numpy.log(1500))))
        N = int(rng.choice([32, 48, 64]))
        nchan = int(rng.integers(1, 4))
        dec = math.radians(rng.uniform(-80, 10))
        a1, _ = simulation.baseline_indices(len(ants), autocorrelations=False)
        ntimes = max(1, nrow // len(a1) + 1)
        ha0 = rng.uniform(-2, 2) * math.pi / 12
        uvw = simulation.baseline_uvw(ants, ha0 + numpy.linspace(0, rng.uniform(2, 20) / 720 * math.pi, ntimes),
                                      dec, autocorrelations=False).reshape(-1, 3)[:nrow]
        freq = 1e8 + 2e6 * numpy.arange(nchan)
        cell = 0.2 / (numpy.abs(uvw[:, :2]).max() * freq.max() / C)
        ms = rng.normal(size=(len(uvw), nchan)) + 1j * rng.normal(size=(len(uvw), nchan))
        img = rng.normal(size=(N, N))
        jobs.append(dict(uvw=uvw, freq=freq, npix=N, cell=cell, nvis=len(uvw) * nchan,
                         ms=torch.tensor(ms, dtype=torch.complex64, device=dev),
                         model=torch.tensor(img, dtype=torch.float32, device=dev)))
        refs.append((ox.explicit_ms2dirty(uvw, freq, ms, None, N, N, cell, cell, True),
                     ox.explicit_dirty2ms(uvw, freq, img, None, cell, cell, True)))
    tuner = AutoTuner()
    batch = Batch(jobs, eps, "mixed", tuner=tuner, nstreams=4)
    assert len(tuner.table) >= 2
    for _ in range(2):
        for d in batch.dirty:
            d.fill_(7.0)
        r = batch.run()
        assert len(r["latency_ms"]) == len(jobs)
        for k, (ref_d, ref_v) in enumerate(refs):
            ed = float(numpy.linalg.norm(batch.dirty[k][0].double().cpu().numpy() - ref_d) / numpy.linalg.norm(ref_d))
            ev = float(numpy.linalg.norm(batch.vis[k][..., 0].cpu().numpy() - ref_v) / numpy.linalg.norm(ref_v))
            assert ed <= eps and ev <= eps, (k, batch.opts[k], ed, ev)
    batch.close()
