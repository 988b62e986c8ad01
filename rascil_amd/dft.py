"""Point-source DFT on the GPU (hipng_dft kernel).

Replaces dft_skycomponent_visibility / extract_direction_and_flux / dft_kernel of
rascil/processing_components/imaging/dft.py:44-178 (CPU loop :261-281, CUDA kernel
:181-258): vis[t, b, c, p] = sum_k flux[k, c, p] exp(-2 pi i (u l_k + v m_k + w (n_k - 1))),
FP64, no 1/n factor. Every dft_compute_kernel choice runs the same HIP kernel.
"""

import collections.abc
import ctypes

import numpy

from . import _lib
from .data_models import BlockVisibility, direction_to_lmn
from .polarisation import convert_pol_frame


def extract_direction_and_flux(sc, vis: BlockVisibility):
    """Direction cosines [K, 3] (l, m, n-1) and fluxes [K, nchan, npol] in the vis frame (dft.py:65-120)."""
    if not isinstance(sc, collections.abc.Iterable):
        sc = [sc]
    vfluxes, dcs = [], []
    for comp in sc:
        flux = numpy.asarray(comp.flux)
        if comp.polarisation_frame != vis.blockvisibility_acc.polarisation_frame:
            flux = convert_pol_frame(flux, comp.polarisation_frame,
                                     vis.blockvisibility_acc.polarisation_frame)
        if not (len(comp.frequency) == len(vis.frequency)
                and numpy.allclose(comp.frequency, vis.frequency, rtol=1e-15)):
            if flux.shape[0] > 1:
                raise ValueError("frequency interpolation of component fluxes is not supported")
        vfluxes.append(numpy.broadcast_to(flux, (len(vis.frequency), flux.shape[-1])))
        l, m, _ = direction_to_lmn(comp.direction, vis.phasecentre)
        dcs.append([l, m, numpy.sqrt(1 - l ** 2 - m ** 2) - 1.0])
    return numpy.array(dcs, dtype="f8"), numpy.array(vfluxes).astype("complex")


def dft_kernel(direction_cosines, vfluxes, uvw_lambda, dft_compute_kernel=None, **kwargs):
    """GPU DFT: direction_cosines [K,3], vfluxes [K, nchan, npol], uvw_lambda [t, b, c, 3]."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("hipng DFT needs a ROCm GPU")
    lib = _lib.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    ntimes, nbl, nchan, _ = uvw_lambda.shape
    ncomp = direction_cosines.shape[0]
    npol = vfluxes.shape[-1]
    dc = torch.from_numpy(numpy.ascontiguousarray(direction_cosines, "f8")).to(dev)
    fl = torch.from_numpy(numpy.ascontiguousarray(
        numpy.broadcast_to(vfluxes, (ncomp, nchan, npol)), "c16")).to(dev)
    uv = torch.from_numpy(numpy.ascontiguousarray(uvw_lambda, "f8")).to(dev)
    out = torch.empty((ntimes, nbl, nchan, npol), dtype=torch.complex128, device=dev)
    _lib.check(lib.hipng_dft(ctypes.c_void_p(dc.data_ptr()), ctypes.c_void_p(fl.data_ptr()),
                             ctypes.c_void_p(uv.data_ptr()), ncomp, ntimes * nbl, nchan, npol,
                             ctypes.c_void_p(out.data_ptr()),
                             ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
    return out.cpu().numpy()


def dft_skycomponent_visibility(vis: BlockVisibility, sc, **kwargs) -> BlockVisibility:
    """Set vis to the DFT of the component(s) (dft.py:44-62)."""
    if sc is None or (isinstance(sc, list) and len(sc) == 0):
        return vis
    direction_cosines, vfluxes = extract_direction_and_flux(sc, vis)
    vis.vis = dft_kernel(direction_cosines, vfluxes, vis.uvw_lambda, **kwargs)
    return vis
