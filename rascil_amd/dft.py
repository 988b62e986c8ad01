"""Point-source DFT and inverse DFT on the GPU (hipng_dft / hipng_idft kernels).

Replaces dft_skycomponent_visibility / extract_direction_and_flux / dft_kernel of
rascil/processing_components/imaging/dft.py:44-178 (CPU loop :261-281, CUDA kernel
:181-258): vis[t, b, c, p] = sum_k flux[k, c, p] exp(-2 pi i (u l_k + v m_k + w (n_k - 1))),
FP64, no 1/n factor; every dft_compute_kernel choice runs the same HIP kernel. And
idft_visibility_skycomponent (:335-380): the weighted mean of vis times the conjugate phasor.
"""

import collections.abc
import copy
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
                # cubic interpolation per polarisation onto the visibility channels (:94-107);
                # like the reference this interpolates the component's own flux (before the
                # polarisation conversion above) into a real array
                from scipy import interpolate

                cflux = numpy.asarray(comp.flux)
                vflux = numpy.zeros([len(vis.frequency), cflux.shape[1]])
                for pol in range(cflux.shape[1]):
                    fint = interpolate.interp1d(comp.frequency, cflux[:, pol], kind="cubic")
                    vflux[:, pol] = fint(vis.frequency)
                flux = vflux
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


def dft_device(direction_cosines, vfluxes, uvw_lambda, out=None):
    """GPU DFT on device tensors (no host copies): direction_cosines [K, 3] float64,
    vfluxes [K, nchan, npol] complex128, uvw_lambda [t, b, nchan, 3] float64 -> vis
    [t, b, nchan, npol] complex128 (or into `out`), on the current stream."""
    import torch

    lib = _lib.load()
    ntimes, nbl, nchan, three = uvw_lambda.shape
    ncomp, fnchan, npol = vfluxes.shape
    dev = uvw_lambda.device
    if three != 3 or tuple(direction_cosines.shape) != (ncomp, 3) or fnchan != nchan:
        raise ValueError("dft_device: shapes [K,3], [K,nchan,npol], [t,b,nchan,3] do not match")
    for name, t, dt in (("direction_cosines", direction_cosines, torch.float64),
                        ("vfluxes", vfluxes, torch.complex128), ("uvw_lambda", uvw_lambda, torch.float64)):
        if t.dtype != dt or not t.is_contiguous() or t.device != dev:
            raise ValueError(f"dft_device: {name} must be a contiguous {dt} tensor on {dev}")
    if out is None:
        out = torch.empty((ntimes, nbl, nchan, npol), dtype=torch.complex128, device=dev)
    elif out.dtype != torch.complex128 or not out.is_contiguous() or tuple(out.shape) != (ntimes, nbl, nchan, npol):
        raise ValueError("dft_device: out must be a contiguous complex128 tensor [t, b, nchan, npol]")
    _lib.check(lib.hipng_dft(ctypes.c_void_p(direction_cosines.data_ptr()), ctypes.c_void_p(vfluxes.data_ptr()),
                             ctypes.c_void_p(uvw_lambda.data_ptr()), ncomp, ntimes * nbl, nchan, npol,
                             ctypes.c_void_p(out.data_ptr()),
                             ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
    return out


def dft_skycomponent_visibility(vis: BlockVisibility, sc, **kwargs) -> BlockVisibility:
    """Set vis to the DFT of the component(s) (dft.py:44-62)."""
    if sc is None or (isinstance(sc, list) and len(sc) == 0):
        return vis
    direction_cosines, vfluxes = extract_direction_and_flux(sc, vis)
    vis.vis = dft_kernel(direction_cosines, vfluxes, vis.uvw_lambda, **kwargs)
    return vis


def idft_visibility_skycomponent(vis: BlockVisibility, sc):
    """Inverse DFT components from the visibilities (dft.py:335-380): for each component the
    flux [nchan, npol] = sum over times and baselines of w (1 - f) vis conj(phasor) divided by
    the sum of w (1 - f) (0 where that is not positive), converted back to the component's
    polarisation frame; returns (components, weights) like the reference. Runs hipng_idft."""
    import torch

    if sc is None:
        return sc
    if not isinstance(sc, collections.abc.Iterable):
        sc = [sc]
    if not torch.cuda.is_available():
        raise RuntimeError("hipng inverse DFT needs a ROCm GPU")
    lib = _lib.load()
    dev = torch.device("cuda", torch.cuda.current_device())
    ntimes, nbl, nchan, npol = vis.vis.shape
    dcs = []
    for comp in sc:
        l, m, _ = direction_to_lmn(comp.direction, vis.phasecentre)
        dcs.append([l, m, numpy.sqrt(1 - l ** 2 - m ** 2) - 1.0])
    ncomp = len(dcs)
    dc = torch.tensor(numpy.array(dcs, "f8").reshape(ncomp, 3), device=dev)
    v = torch.from_numpy(numpy.ascontiguousarray(vis.vis, "c16")).to(dev)
    w = torch.from_numpy(numpy.ascontiguousarray(vis.weight, "f8")).to(dev)
    f = torch.from_numpy(numpy.ascontiguousarray(vis.flags, numpy.int32)).to(dev)
    uv = torch.from_numpy(numpy.ascontiguousarray(vis.uvw_lambda, "f8")).to(dev)
    flux = torch.empty((ncomp, nchan, npol), dtype=torch.complex128, device=dev)
    wsum = torch.empty((nchan, npol), dtype=torch.float64, device=dev)
    _lib.check(lib.hipng_idft(ctypes.c_void_p(dc.data_ptr()), ctypes.c_void_p(v.data_ptr()),
                              ctypes.c_void_p(w.data_ptr()), ctypes.c_void_p(f.data_ptr()),
                              ctypes.c_void_p(uv.data_ptr()), ncomp, ntimes * nbl, nchan, npol,
                              ctypes.c_void_p(flux.data_ptr()), ctypes.c_void_p(wsum.data_ptr()),
                              ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
    flux = flux.cpu().numpy()
    weight = wsum.cpu().numpy()
    newsc, weights = [], []
    for k, comp in enumerate(sc):
        newcomp = copy.deepcopy(comp)
        fk = flux[k]
        if comp.polarisation_frame != vis.blockvisibility_acc.polarisation_frame:
            fk = convert_pol_frame(fk, vis.blockvisibility_acc.polarisation_frame, comp.polarisation_frame)
        newcomp.flux = fk
        newsc.append(newcomp)
        weights.append(weight)
    return newsc, weights
