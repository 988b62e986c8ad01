"""Differentiable predict: the degridder as a torch.autograd.Function whose backward is the
gridder (its exact adjoint) and whose double backward is the degridder again.

For a real image I and the linear operator A = dirty2ms (vis = A I), a real loss L gives
dL/dI = Re(A^H dL/dV*) and Re(A^H g) is exactly ms2dirty(g) without weights. Gains
V = g_a1 conj(g_a2) A I (rascil/processing_components/calibration/operations.py:135-137
convention) are ordinary torch complex ops around it, so image and gain gradients of a
visibility chi-squared come from the GPU without forming the measurement matrix.
"""

import torch

from . import wgridder


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


def predict(image, uvw=None, freq=None, pixsize_x=None, pixsize_y=None, epsilon=1e-7,
            do_wstacking=True, precision=None, geometry=None):
    """Differentiable dirty2ms.

    image: real GPU tensor [nx, ny] (or [nb, nx, ny]); returns complex vis [nrow, nchan] (or
    [nrow, nchan, nb]). Pass a prebuilt wgridder.Geometry to reuse its bucket sort across
    calls (forward, backward and double backward all use it); otherwise one is built from
    uvw [nrow, 3] (metres, ducc0 sign convention), freq (Hz) and the pixel sizes.
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
    return _Degrid.apply(image, geometry)


def apply_gains(vis, gains, ant1, ant2):
    """V_ij = g_i conj(g_j) M_ij for vis [nrow, nchan(, nb)] with per-row antenna indices."""
    g = gains[ant1] * torch.conj(gains[ant2])
    while g.dim() < vis.dim():
        g = g.unsqueeze(-1)
    return g * vis
