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


class _Geometry:
    def __init__(self, uvw, freq, npix_x, npix_y, pixsize_x, pixsize_y, epsilon, do_wstacking,
                 precision):
        self.uvw = uvw
        self.freq = freq
        self.npix_x, self.npix_y = npix_x, npix_y
        self.pixsize_x, self.pixsize_y = pixsize_x, pixsize_y
        self.epsilon = epsilon
        self.do_wstacking = do_wstacking
        self.precision = precision

    def degrid(self, image):
        return wgridder.dirty2ms(self.uvw, self.freq, image.contiguous(), None, self.pixsize_x,
                                 self.pixsize_y, 0, 0, self.epsilon, self.do_wstacking,
                                 precision=self.precision)

    def grid(self, vis):
        return wgridder.ms2dirty(self.uvw, self.freq, vis.contiguous(), None, self.npix_x,
                                 self.npix_y, self.pixsize_x, self.pixsize_y, 0, 0, self.epsilon,
                                 self.do_wstacking, precision=self.precision)


class _Degrid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, geom):
        ctx.geom = geom
        return geom.degrid(image)

    @staticmethod
    def backward(ctx, grad_vis):
        return _Grid.apply(grad_vis, ctx.geom), None


class _Grid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vis, geom):
        ctx.geom = geom
        return geom.grid(vis)

    @staticmethod
    def backward(ctx, grad_image):
        # adjoint of Re(A^H v) as a map C -> R: the real image gradient goes back through A
        return _Degrid.apply(grad_image.to(grad_image.dtype), ctx.geom), None


def predict(image, uvw, freq, pixsize_x, pixsize_y=None, epsilon=1e-7, do_wstacking=True,
            precision=None):
    """Differentiable dirty2ms: image [nx, ny] (or [nb, nx, ny]) real GPU tensor -> vis
    [nrow, nchan] (or [nrow, nchan, nb]) complex GPU tensor. uvw is a [nrow, 3] FP64 GPU
    tensor in metres with ducc0's sign convention; freq a host array in Hz."""
    if pixsize_y is None:
        pixsize_y = pixsize_x
    if precision is None:
        precision = "fp64" if image.dtype == torch.float64 else "mixed"
    nx, ny = image.shape[-2], image.shape[-1]
    geom = _Geometry(uvw, freq, nx, ny, pixsize_x, pixsize_y, epsilon, do_wstacking, precision)
    return _Degrid.apply(image, geom)


def apply_gains(vis, gains, ant1, ant2):
    """V_ij = g_i conj(g_j) M_ij for vis [nrow, nchan(, nb)] with per-row antenna indices."""
    g = gains[ant1] * torch.conj(gains[ant2])
    while g.dim() < vis.dim():
        g = g.unsqueeze(-1)
    return g * vis
