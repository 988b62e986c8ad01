"""Differentiable predict: gradients of a visibility chi-squared with respect to the image
and per-antenna complex gains, checked by central finite differences (FP64 1e-6, FP32 1e-3)
and against the adjoint of the FP64 brute-force sum (oracle)."""

import math

import numpy
import pytest

from oracle import explicit as ox
from rascil_amd import Direction, simulation
from rascil_amd.autograd import apply_gains, predict

pytestmark = pytest.mark.gpu


def _setup(dtype, nant=12, ntimes=4, npix=32, cell=2e-3, seed=0):
    import torch

    rng = numpy.random.default_rng(seed)
    ants = simulation.gaussian_disc_stations(nant, 300.0, 1000.0, rng)
    uvw = simulation.baseline_uvw(ants, numpy.linspace(-0.5, 0.5, ntimes) * math.pi / 12,
                                  math.radians(-45), autocorrelations=False).reshape(-1, 3)
    a1, a2 = simulation.baseline_indices(nant, autocorrelations=False)
    ant1 = numpy.tile(a1, ntimes)
    ant2 = numpy.tile(a2, ntimes)
    freq = numpy.array([1.0e8, 1.2e8])
    dev = torch.device("cuda", 0)
    img = torch.tensor(rng.normal(size=(npix, npix)), dtype=dtype, device=dev, requires_grad=True)
    cdt = torch.complex128 if dtype == torch.float64 else torch.complex64
    gains = torch.tensor(rng.uniform(0.9, 1.1, nant) * numpy.exp(1j * rng.uniform(-0.5, 0.5, nant)),
                         dtype=cdt, device=dev, requires_grad=True)
    vobs = torch.tensor(rng.normal(size=(uvw.shape[0], 2)) + 1j * rng.normal(size=(uvw.shape[0], 2)),
                        dtype=cdt, device=dev)
    uvw_d = torch.tensor(uvw, dtype=torch.float64, device=dev)
    i1 = torch.tensor(ant1, device=dev)
    i2 = torch.tensor(ant2, device=dev)
    return dict(uvw=uvw, uvw_d=uvw_d, freq=freq, img=img, gains=gains, vobs=vobs, i1=i1, i2=i2,
                cell=cell)


def _loss(s, img, gains, eps, prec):
    import torch

    v = predict(img, s["uvw_d"], s["freq"], s["cell"], epsilon=eps, do_wstacking=True, precision=prec)
    # chi-squared accumulated in FP64 (the FP32 path's gradients flow back through the casts)
    r = s["vobs"].to(torch.complex128) - apply_gains(v.to(torch.complex128), gains.to(torch.complex128),
                                                     s["i1"], s["i2"])
    return (r.abs() ** 2).sum()


# The loss is exactly quadratic in the image, so the central difference is exact for any step:
# a large image step keeps the FP32 forward error out of the difference quotient.
@pytest.mark.parametrize("prec,eps,tol,h,hg", [("fp64", 1e-12, 1e-6, 1e-3, 1e-5),
                                               ("mixed", 1e-5, 1e-3, 1.0, 1e-2)])
def test_image_and_gain_gradients_finite_differences(prec, eps, tol, h, hg):
    import torch

    dtype = torch.float64 if prec == "fp64" else torch.float32
    s = _setup(dtype)
    L = _loss(s, s["img"], s["gains"], eps, prec)
    L.backward()
    gi = s["img"].grad.detach().cpu().numpy()
    gg = s["gains"].grad.detach().cpu().numpy()
    rng = numpy.random.default_rng(9)
    with torch.no_grad():
        for _ in range(5):
            ix, iy = rng.integers(0, 32, 2)
            ip = s["img"].detach().clone()
            im = s["img"].detach().clone()
            ip[ix, iy] += h
            im[ix, iy] -= h
            fd = (_loss(s, ip, s["gains"], eps, prec) - _loss(s, im, s["gains"], eps, prec)).item() / (2 * h)
            assert abs(fd - gi[ix, iy]) <= tol * max(1.0, abs(fd)), (fd, gi[ix, iy])
        for k in (0, 5):
            for d in (1.0, 1j):  # torch grad = dL/dRe + i dL/dIm
                gp = s["gains"].detach().clone()
                gm = s["gains"].detach().clone()
                gp[k] += hg * d
                gm[k] -= hg * d
                fd = (_loss(s, s["img"], gp, eps, prec) - _loss(s, s["img"], gm, eps, prec)).item() / (2 * hg)
                an = gg[k].real if d == 1.0 else gg[k].imag
                assert abs(fd - an) <= tol * max(1.0, abs(fd)), (k, d, fd, an)


def test_image_gradient_matches_dft_adjoint():
    """dL/dI = -2 Re(A^H (g_i conj g_j)^* r) with A the exact FP64 sum (oracle)."""
    import torch

    s = _setup(torch.float64)
    L = _loss(s, s["img"], s["gains"], 1e-12, "fp64")
    L.backward()
    with torch.no_grad():
        v = predict(s["img"], s["uvw_d"], s["freq"], s["cell"], epsilon=1e-12)
        g = s["gains"][s["i1"]] * torch.conj(s["gains"][s["i2"]])
        r = s["vobs"] - g[:, None] * v
        back = (-2.0 * torch.conj(g)[:, None] * r).cpu().numpy()
    ref = ox.explicit_ms2dirty(s["uvw"], s["freq"], back, None, 32, 32, s["cell"], s["cell"], True)
    gi = s["img"].grad.cpu().numpy()
    assert numpy.linalg.norm(gi - ref) / numpy.linalg.norm(ref) < 1e-10


def test_double_backward_is_degridder():
    """Hessian-vector product of 0.5|A I|^2 equals A^H A v, computed through double backward."""
    import torch

    s = _setup(torch.float64)
    img = s["img"]
    v = predict(img, s["uvw_d"], s["freq"], s["cell"], epsilon=1e-12)
    L = 0.5 * (v.abs() ** 2).sum()
    (g,) = torch.autograd.grad(L, img, create_graph=True)
    vec = torch.randn_like(img)
    (hv,) = torch.autograd.grad((g * vec).sum(), img)
    with torch.no_grad():
        av = predict(vec, s["uvw_d"], s["freq"], s["cell"], epsilon=1e-12)
        from rascil_amd import wgridder

        ref = wgridder.ms2dirty(s["uvw_d"], s["freq"], av, None, 32, 32, s["cell"], s["cell"], 0, 0,
                                1e-12, True)
    assert torch.linalg.norm(hv - ref) / torch.linalg.norm(ref) < 1e-10


def test_gain_descent_reduces_chi_squared():
    import torch

    s = _setup(torch.float64)
    with torch.no_grad():
        truth = s["gains"].detach().clone()
        v = predict(s["img"], s["uvw_d"], s["freq"], s["cell"], epsilon=1e-10)
        s["vobs"] = apply_gains(v, truth, s["i1"], s["i2"])
    g = torch.ones_like(truth).requires_grad_(True)
    opt = torch.optim.Adam([g], lr=0.02)
    first = None
    for it in range(50):
        opt.zero_grad()
        L = _loss(s, s["img"].detach(), g, 1e-10, "fp64")
        L.backward()
        opt.step()
        first = first if first is not None else L.item()
    assert L.item() < 0.2 * first


@pytest.mark.parametrize("prec,eps,tol", [("fp64", 1e-12, 1e-12), ("mixed", 1e-5, 1e-5)])
def test_fused_gains_match_torch_gains_forward_and_gradients(prec, eps, tol):
    """predict(..., gains=) (gains fused into the degridder's unsort, the conjugate product
    into the gridder's value gather in backward, the gain gradient from hipng_gain_gradient)
    gives the same visibilities and the same image and gain gradients as predict followed by
    torch's elementwise apply_gains (calibration/operations.py:135-137 convention)."""
    import torch

    from rascil_amd import wgridder

    dtype = torch.float64 if prec == "fp64" else torch.float32
    s = _setup(dtype)
    geom = wgridder.Geometry(s["uvw_d"], s["freq"], 32, 32, s["cell"], s["cell"], eps, True, prec)
    i1, i2 = s["i1"].to(torch.int32), s["i2"].to(torch.int32)

    def loss_fused(img, g):
        v = predict(img, geometry=geom, gains=g, ant1=i1, ant2=i2).to(torch.complex128)
        return ((s["vobs"].to(torch.complex128) - v).abs() ** 2).sum(), v

    def loss_torch(img, g):
        v = apply_gains(predict(img, geometry=geom).to(torch.complex128), g.to(torch.complex128), s["i1"], s["i2"])
        return ((s["vobs"].to(torch.complex128) - v).abs() ** 2).sum(), v

    img_a, g_a = s["img"].detach().clone().requires_grad_(True), s["gains"].detach().clone().requires_grad_(True)
    img_b, g_b = s["img"].detach().clone().requires_grad_(True), s["gains"].detach().clone().requires_grad_(True)
    La, va = loss_fused(img_a, g_a)
    Lb, vb = loss_torch(img_b, g_b)
    assert (torch.linalg.norm(va - vb) / torch.linalg.norm(vb)).item() < tol
    La.backward()
    Lb.backward()
    ei = (torch.linalg.norm(img_a.grad - img_b.grad) / torch.linalg.norm(img_b.grad)).item()
    eg = (torch.linalg.norm(g_a.grad.to(torch.complex128) - g_b.grad.to(torch.complex128))
          / torch.linalg.norm(g_b.grad.to(torch.complex128))).item()
    print(f"{prec}: fused vs torch gains: image grad {ei:.2e}, gain grad {eg:.2e}")
    assert ei < 10 * tol and eg < 10 * tol
    geom.close()


def test_fused_gains_gradients_finite_differences():
    """Central differences of the fused-gain chi-squared in FP64: image (exact, the loss is
    quadratic) and the real and imaginary parts of two antennas' gains, within 1e-6."""
    import torch

    from rascil_amd import wgridder

    s = _setup(torch.float64)
    geom = wgridder.Geometry(s["uvw_d"], s["freq"], 32, 32, s["cell"], s["cell"], 1e-12, True, "fp64")
    i1, i2 = s["i1"].to(torch.int32), s["i2"].to(torch.int32)

    def loss(img, g):
        v = predict(img, geometry=geom, gains=g, ant1=i1, ant2=i2)
        return ((s["vobs"] - v).abs() ** 2).sum()

    img0, g0 = s["img"].detach().clone(), s["gains"].detach().clone()
    img, g = img0.clone().requires_grad_(True), g0.clone().requires_grad_(True)
    loss(img, g).backward()
    with torch.no_grad():
        for px, py in ((3, 7), (16, 16)):
            ip, im_ = img0.clone(), img0.clone()
            ip[px, py] += 1e-3
            im_[px, py] -= 1e-3
            fd = (loss(ip, g0) - loss(im_, g0)).item() / 2e-3
            assert abs(fd - img.grad[px, py].item()) <= 1e-6 * max(1.0, abs(fd))
        for k in (0, 5):
            for dirn in (1.0, 1j):
                gp, gm = g0.clone(), g0.clone()
                gp[k] += 1e-5 * dirn
                gm[k] -= 1e-5 * dirn
                fd = (loss(img0, gp) - loss(img0, gm)).item() / 2e-5
                an = g.grad[k].real.item() if dirn == 1.0 else g.grad[k].imag.item()
                assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (k, dirn, fd, an)
    geom.close()
