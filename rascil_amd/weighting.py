"""Imaging weights on the GPU: weight_visibility (natural / uniform / robust).

Mirrors rascil/processing_components/imaging/weighting.py:28-52, which grids the flagged
imaging weights onto an image-sized uv grid (griddata/gridding.py:260-308, nearest cell
via the grid WCS of griddata/operations.py:71-107) and reweights every visibility from the
density of its cell (griddata/gridding.py:336-424). The density grid and reweighting run
in hipng_weight_visibility.
"""

import ctypes

import numpy

from . import _lib
from .data_models import image_is_canonical


def weight_visibility(vis, model, weighting="uniform", robustness=0.0, **kwargs):
    """Set vis.imaging_weight for natural, uniform or robust weighting; returns vis."""
    import torch

    assert image_is_canonical(model)
    assert weighting in ["natural", "uniform", "robust"], "Weighting {} not supported".format(weighting)
    if weighting == "natural":
        vis.imaging_weight = vis.weight.copy()
        return vis
    if not torch.cuda.is_available():
        raise RuntimeError("hipng weighting needs a ROCm GPU")
    lib = _lib.load()
    nimchan, _, ny, nx = model.pixels.shape
    d2r = numpy.pi / 180.0
    cell_u = 1.0 / (nx * model.wcs.cdelt[0] * d2r)
    cell_v = 1.0 / (ny * model.wcs.cdelt[1] * d2r)
    vis_to_im = numpy.ascontiguousarray(
        numpy.round(model.image_acc.wcs.freq_to_pix(vis.frequency)).astype(numpy.int32))
    nt, nb, nchan, npol = vis.vis.shape
    dev = torch.device("cuda", torch.cuda.current_device())
    uvw = torch.from_numpy(numpy.ascontiguousarray(vis.uvw.reshape(-1, 3), "f8")).to(dev)
    wt = torch.from_numpy(numpy.ascontiguousarray(
        (vis.imaging_weight * (1 - vis.flags)).reshape(-1, nchan, npol), "f8")).to(dev)
    grid = torch.empty((nimchan, ny, nx, npol), dtype=torch.float64, device=dev)
    freq = numpy.ascontiguousarray(vis.frequency, "f8")
    _lib.check(lib.hipng_weight_visibility(
        ctypes.c_void_p(uvw.data_ptr()), _lib.dptr(freq), nt, nb, nchan, npol,
        vis_to_im.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), nimchan, nx, ny, cell_u, cell_v,
        1 if weighting == "uniform" else 2, float(robustness), ctypes.c_void_p(wt.data_ptr()),
        ctypes.c_void_p(grid.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
    vis.imaging_weight = wt.cpu().numpy().reshape(nt, nb, nchan, npol)
    return vis
