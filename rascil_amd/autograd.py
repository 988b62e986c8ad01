"""Differentiable predict: the degridder as a torch.autograd.Function whose backward is the
gridder (its exact adjoint) and whose double backward is the degridder again.

For a real image I and the linear operator A = dirty2ms (vis = A I), a real loss L gives
dL/dI = Re(A^H dL/dV*) and Re(A^H g) is exactly ms2dirty(g) without weights. Per-antenna
gains V = g_a1 conj(g_a2) A I (rascil/processing_components/calibration/operations.py:135-137
convention) are fused into the degridder's unsort (forward) and the gridder's value gather
(backward: conj(g_a1 conj(g_a2)) dL/dV), and the gain gradient is one reduction kernel over
the saved model visibilities (hipng_gain_gradient), so image and gain gradients of a
visibility chi-squared come from the GPU without forming the measurement matrix and without
extra passes over the visibilities.
"""

import numpy
import torch

from . import _lib, wgridder


class _Degrid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, geom):
        ctx.geom = geom
        ctx.image_shape = image.shape
        vis = geom.dirty2ms(image.contiguous())
        return vis.reshape(geom.nrow, geom.nchan) if image.dim() == 2 else vis

    @staticmethod
    def backward(ctx, grad_vis):
        return _Grid.apply(grad_vis, ctx.geom, ctx.image_shape), None


class _Grid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vis, geom, image_shape):
        ctx.geom = geom
        ctx.vis_shape = vis.shape
        return geom.ms2dirty(vis.contiguous()).reshape(image_shape)

    @staticmethod
    def backward(ctx, grad_image):
        # Re(A^H v) maps C -> R; its adjoint sends the real image gradient back through A
        return _Degrid.apply(grad_image, ctx.geom).reshape(ctx.vis_shape), None, None


def _gain_columns(geom, gains128, ant1, ant2):
    return wgridder.make_columns(geom.nchan, geom.nbatch, numpy.eye(geom.nbatch), vis_dtype=geom.cdtype,
                                 gains=gains128, ant1=ant1, ant2=ant2, image_order=0, image_dtype=geom.rdtype)


def _expand_gains(gains, nchan):
    """Per-antenna gains [nant] (or per antenna and channel [nant, nchan]) as complex128 [nant, nchan]."""
    g = gains.to(torch.complex128)
    if g.dim() == 1:
        g = g[:, None].expand(g.shape[0], nchan)
    return g.contiguous()


class _DegridGains(torch.autograd.Function):
    """V = g[ant1] conj(g[ant2]) A I in one fused degridder pass; backward: the gridder with
    the conjugate gain product fused into its value gather (image gradient) and the gain
    gradient kernel over the model visibilities A I saved by the forward pass."""

    @staticmethod
    def forward(ctx, image, gains, geom, ant1, ant2):
        g128 = _expand_gains(gains, geom.nchan)
        cols = _gain_columns(geom, g128, ant1, ant2)
        img = image.reshape(geom.nbatch, geom.npix_x, geom.npix_y).contiguous()
        vis = torch.empty((geom.nrow, geom.nchan, geom.nbatch), dtype=geom.cdtype, device=geom.device)
        model = torch.empty_like(vis) if gains.requires_grad else None
        geom.predict_columns(cols, img, vis, model)
        ctx.geom, ctx.ant1, ctx.ant2 = geom, ant1, ant2
        ctx.image_shape, ctx.gains_shape, ctx.gains_dtype = image.shape, gains.shape, gains.dtype
        ctx.save_for_backward(g128, model if model is not None else torch.empty(0, device=geom.device))
        return vis.reshape(geom.nrow, geom.nchan) if image.dim() == 2 else vis

    @staticmethod
    def backward(ctx, grad_vis):
        geom = ctx.geom
        g128, model = ctx.saved_tensors
        gv = grad_vis.reshape(geom.nrow, geom.nchan, geom.nbatch).to(geom.cdtype).contiguous()
        grad_image = grad_gains = None
        if ctx.needs_input_grad[0]:
            cols = _gain_columns(geom, g128, ctx.ant1, ctx.ant2)
            grad_image = geom.invert_columns(cols, gv).reshape(ctx.image_shape)
        if ctx.needs_input_grad[1]:
            nant = g128.shape[0]
            gg = torch.empty((nant, geom.nchan), dtype=torch.complex128, device=geom.device)
            vt = _lib.HIPNG_COMPLEX64 if geom.cdtype == torch.complex64 else _lib.HIPNG_COMPLEX128
            vp = wgridder.ctypes.c_void_p
            _lib.check(_lib.load().hipng_gain_gradient(
                vp(model.data_ptr()), vp(gv.data_ptr()), vt, vp(g128.data_ptr()), vp(ctx.ant1.data_ptr()),
                vp(ctx.ant2.data_ptr()), geom.nrow, geom.nchan, geom.nbatch, nant, vp(gg.data_ptr()),
                wgridder._stream(torch, geom.device)))
            grad_gains = (gg.sum(dim=1) if len(ctx.gains_shape) == 1 else gg).to(ctx.gains_dtype)
        return grad_image, grad_gains, None, None, None


def predict(image, uvw=None, freq=None, pixsize_x=None, pixsize_y=None, epsilon=1e-7,
            do_wstacking=True, precision=None, geometry=None, gains=None, ant1=None, ant2=None):
    """Differentiable dirty2ms.

    image: real GPU tensor [nx, ny] (or [nb, nx, ny]); returns complex vis [nrow, nchan] (or
    [nrow, nchan, nb]). Pass a prebuilt wgridder.Geometry to reuse its bucket sort across
    calls (forward, backward and double backward all use it); otherwise one is built from
    uvw [nrow, 3] (metres, ducc0 sign convention), freq (Hz) and the pixel sizes.
    gains (complex [nant] per antenna, or [nant, nchan]) with ant1 / ant2 (int32 [nrow]):
    returns g[ant1] conj(g[ant2]) A image with the gains fused into the degridder, and
    gradients with respect to both the image and the gains (no double backward).
    """
    if geometry is None:
        if pixsize_y is None:
            pixsize_y = pixsize_x
        if precision is None:
            precision = "fp64" if image.dtype == torch.float64 else "mixed"
        nb = 1 if image.dim() == 2 else image.shape[0]
        geometry = wgridder.Geometry(uvw, freq, image.shape[-2], image.shape[-1], pixsize_x,
                                     pixsize_y, epsilon, do_wstacking, precision, nb,
                                     device=image.device)
    if gains is not None:
        if ant1 is None or ant2 is None:
            raise ValueError("gains need ant1 and ant2 (int32 antenna indices per row)")
        return _DegridGains.apply(image, gains, geometry, ant1.to(torch.int32).contiguous(),
                                  ant2.to(torch.int32).contiguous())
    return _Degrid.apply(image, geometry)


def apply_gains(vis, gains, ant1, ant2):
    """V_ij = g_i conj(g_j) M_ij for vis [nrow, nchan(, nb)] with per-row antenna indices."""
    g = gains[ant1] * torch.conj(gains[ant2])
    while g.dim() < vis.dim():
        g = g.unsqueeze(-1)
    return g * vis
