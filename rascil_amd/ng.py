"""invert_ng / predict_ng on the hipng GPU gridder.

Drop-in for rascil/processing_components/imaging/ng.py (predict_ng :33-133,
invert_ng :136-282): same signatures, keyword parameters (threads, epsilon,
do_wstacking, verbosity), conventions (u and w negated before the gridder, images
transposed to [x, y], MFS when the image has one channel, per-channel otherwise via the
WCS frequency axis) and results. The ducc0.wgridder calls are replaced by
rascil_amd.wgridder, which runs on the GPU and processes all polarisations of a call in
one batched pass instead of one call per polarisation. The visibility passes around the
gridder (phase shift to the image centre, flag mask, polarisation conversion, precision
and layout) run as one fused GPU pass each way (hipng_vis_prologue / hipng_vis_epilogue).

Extra keyword: precision = "fp64" (default; the reference computes in FP64 with FP64
accumulation), "mixed" or "fp32".
"""

import ctypes
import logging

import numpy

from . import _lib, wgridder
from .data_models import (BlockVisibility, Image, copy_visibility, direction_to_lmn, get_parameter,
                          image_is_canonical)
from .imaging_base import normalise_sumwt
from .polarisation import pol_matrix

log = logging.getLogger("rascil-logger")

__all__ = ["predict_ng", "invert_ng"]


def _pixsize(im):
    return numpy.abs(numpy.radians(im.image_acc.wcs.cdelt[0]))


def _vis_to_im(model, freq):
    return numpy.round(model.image_acc.wcs.freq_to_pix(freq)).astype("int")


def _dtypes(precision):
    return ("c16", "f8") if precision == "fp64" else ("c8", "f4")


def _shift_lmn(vis_pc, im_pc):
    """(l, m, n-1) of the image centre for shift_vis_to_image (imaging/base.py:67-106 with
    phaserotate_visibility visibility/base.py:296-357): None when the centres coincide to
    1e-15 rad or the rotation is the identity (|n-1| < 1e-15); `moved` says whether the
    visibility phase centre becomes the image's."""
    moved = vis_pc.separation(im_pc) > 1e-15
    if not moved:
        return None, False
    l, m, nm1 = direction_to_lmn(im_pc, vis_pc)
    if numpy.abs(nm1) < 1e-15:
        return None, True
    return numpy.array([l, m, nm1], dtype="f8"), True


def _conv(ipf, opf):
    """Conversion matrix [opf.npol][ipf.npol] interleaved (re, im) for the fused passes."""
    cm = numpy.ascontiguousarray(pol_matrix(ipf, opf), dtype=complex)
    return numpy.ascontiguousarray(cm.view(numpy.float64).ravel()), cm.shape


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def predict_ng(bvis: BlockVisibility, model: Image, **kwargs) -> BlockVisibility:
    """Predict visibilities from a model image (reference ng.py:33-133).

    The degridder's output goes through one fused GPU pass (hipng_vis_epilogue): the
    polarisation conversion to the visibility frame and the phase shift of
    shift_vis_to_image(inverse=True).

    :param bvis: BlockVisibility to be predicted (not modified)
    :param model: model image
    :return: new BlockVisibility with the predicted vis
    """
    assert image_is_canonical(model)
    if model is None:
        return bvis
    torch = wgridder._torch()
    lib = _lib.load()
    epsilon = get_parameter(kwargs, "epsilon", 1e-12)
    do_wstacking = get_parameter(kwargs, "do_wstacking", True)
    verbosity = get_parameter(kwargs, "verbosity", 0)
    precision = get_parameter(kwargs, "precision", "fp64")
    cdt, rdt = _dtypes(precision)

    newbvis = copy_visibility(bvis, zero=True)
    freq = numpy.ascontiguousarray(bvis.frequency, dtype="f8")
    nrows, nbaselines, vnchan, vnpol = bvis.vis.shape
    m_nchan, m_npol, ny, nx = model.pixels.shape
    assert m_npol == vnpol
    conv, _ = _conv(model.image_acc.polarisation_frame, bvis.blockvisibility_acc.polarisation_frame)
    lmn, moved = _shift_lmn(newbvis.phasecentre, model.phasecentre)

    dev = torch.device("cuda", torch.cuda.current_device())
    uvw = newbvis.uvw.reshape([nrows * nbaselines, 3])
    fuvw = numpy.nan_to_num(uvw) * numpy.array([-1.0, 1.0, -1.0])
    fuvw_d = torch.from_numpy(numpy.ascontiguousarray(fuvw)).to(dev)
    pixsize = _pixsize(model)
    vis_to_im = _vis_to_im(model, freq)
    tdt = torch.complex128 if precision == "fp64" else torch.complex64
    if m_nchan == 1:
        dirty = torch.from_numpy(numpy.ascontiguousarray(model.pixels[0].transpose(0, 2, 1), dtype=rdt)).to(dev)
        vist = wgridder.dirty2ms(fuvw_d, freq, dirty, None, pixsize, pixsize, 0, 0, epsilon, do_wstacking,
                                 precision=precision, verbosity=verbosity).reshape(nrows * nbaselines, vnchan, vnpol)
    else:
        vist = torch.empty((nrows * nbaselines, vnchan, vnpol), dtype=tdt, device=dev)
        for vchan in range(vnchan):
            imchan = vis_to_im[vchan]
            dirty = torch.from_numpy(
                numpy.ascontiguousarray(model.pixels[imchan].transpose(0, 2, 1), dtype=rdt)).to(dev)
            vist[:, vchan : vchan + 1, :] = wgridder.dirty2ms(
                fuvw_d, freq[vchan : vchan + 1], dirty, None, pixsize, pixsize, 0, 0, epsilon,
                do_wstacking, precision=precision, verbosity=verbosity).reshape(nrows * nbaselines, 1, vnpol)
    vist = vist.contiguous()
    uvw_d = torch.from_numpy(numpy.ascontiguousarray(uvw, dtype="f8")).to(dev) if lmn is not None else None
    out = torch.empty((nrows * nbaselines, vnchan, vnpol), dtype=torch.complex128, device=dev)
    _lib.check(lib.hipng_vis_epilogue(
        ctypes.c_void_p(vist.data_ptr()), ctypes.c_void_p(uvw_d.data_ptr() if uvw_d is not None else None),
        _ptr(freq), nrows * nbaselines, vnchan, m_npol, vnpol, _ptr(conv), _ptr(lmn), _lib.PRECISIONS[precision],
        ctypes.c_void_p(out.data_ptr()), wgridder._stream(torch, dev)))
    newbvis.vis = out.cpu().numpy().reshape([nrows, nbaselines, vnchan, vnpol])
    if moved:
        newbvis.phasecentre = model.phasecentre
        newbvis.attrs["phasecentre"] = model.phasecentre
    return newbvis


def invert_ng(bvis: BlockVisibility, model: Image, dopsf: bool = False, normalise: bool = True,
              **kwargs):
    """Dirty image (or PSF) and sum of weights (reference ng.py:136-282).

    The visibility passes before the gridder (phase shift to the image centre, flag mask,
    polarisation conversion, the gridder's precision and layout) run as one fused GPU pass
    (hipng_vis_prologue); the input is not modified.

    :param bvis: BlockVisibility to be inverted
    :param model: image template (not changed)
    :param dopsf: make the PSF instead of the dirty image
    :param normalise: normalise by the sum of weights
    :return: (image, sumwt [nchan, npol])
    """
    assert image_is_canonical(model)
    torch = wgridder._torch()
    lib = _lib.load()
    im = model.copy(deep=True)
    epsilon = get_parameter(kwargs, "epsilon", 1e-12)
    do_wstacking = get_parameter(kwargs, "do_wstacking", True)
    verbosity = get_parameter(kwargs, "verbosity", 0)
    precision = get_parameter(kwargs, "precision", "fp64")
    cdt, rdt = _dtypes(precision)

    freq = numpy.ascontiguousarray(bvis.frequency, dtype="f8")
    nrows, nbaselines, vnchan, vnpol = bvis.vis.shape
    nchan, npol, ny, nx = im.pixels.shape
    conv, cshape = _conv(bvis.blockvisibility_acc.polarisation_frame, im.image_acc.polarisation_frame)
    if cshape[0] != npol:
        raise ValueError("image polarisation frame does not match the image")
    if npol > vnpol:
        # the reference indexes the weights of the visibility frame by image polarisation
        raise IndexError(f"index {vnpol} is out of bounds for axis 2 with size {vnpol}")
    lmn, _ = _shift_lmn(bvis.phasecentre, im.phasecentre)

    n = nrows * nbaselines
    dev = torch.device("cuda", torch.cuda.current_device())
    st = wgridder._stream(torch, dev)
    uvw = numpy.ascontiguousarray(bvis.uvw.reshape([n, 3]), dtype="f8")
    uvw_d = torch.from_numpy(uvw).to(dev)
    vis_d = torch.from_numpy(numpy.ascontiguousarray(bvis.vis, dtype="c16")).to(dev)
    flags = numpy.asarray(bvis.flags)
    flags_d = torch.from_numpy(numpy.ascontiguousarray(flags, dtype=numpy.int32)).to(dev) if flags.any() else None
    iwt_d = torch.from_numpy(numpy.ascontiguousarray(bvis.imaging_weight, dtype="f8")).to(dev)
    tdt = torch.complex128 if precision == "fp64" else torch.complex64
    wdt = torch.float64 if precision == "fp64" else torch.float32
    ms = torch.empty((n, vnchan, npol), dtype=tdt, device=dev)
    wgt = torch.empty((n, vnchan, npol), dtype=wdt, device=dev)
    nz = torch.empty((vnchan, npol), dtype=torch.int32, device=dev)
    _lib.check(lib.hipng_vis_prologue(
        ctypes.c_void_p(vis_d.data_ptr()), ctypes.c_void_p(flags_d.data_ptr() if flags_d is not None else None),
        ctypes.c_void_p(iwt_d.data_ptr()), ctypes.c_void_p(uvw_d.data_ptr()), _ptr(freq), n, vnchan, vnpol, npol,
        _ptr(conv), _ptr(lmn), _lib.PRECISIONS[precision], ctypes.c_void_p(ms.data_ptr()),
        ctypes.c_void_p(wgt.data_ptr()), ctypes.c_void_p(nz.data_ptr()), st))
    del vis_d, iwt_d, flags_d
    nonzero = nz.cpu().numpy() != 0
    # sum of the flagged imaging weights per (channel, pol) of the visibility frame, in FP64
    wsum = numpy.sum(bvis.imaging_weight * (1 - flags), axis=(0, 1))

    npixdirty = im.pixels.shape[-1]
    pixsize = _pixsize(im)
    fuvw_d = uvw_d * torch.tensor([-1.0, 1.0, -1.0], dtype=torch.float64, device=dev)
    im.pixels[...] = 0.0
    sumwt = numpy.zeros([nchan, npol])
    vis_to_im = _vis_to_im(model, freq)
    mfs = nchan == 1 and vnchan > 1
    if dopsf:
        ms.zero_()
        ms[..., 0] = 1.0
        nonzero[...] = False
        nonzero[:, 0] = True

    def grid(sel_ms, sel_wgt, sel_freq, active_mask):
        # polarisations with data are gridded together; all-zero ones are skipped as in the
        # reference (their weights still enter sumwt)
        active = [p for p in range(npol) if active_mask[p]]
        out = {}
        if active:
            if len(active) == npol:
                lms, lwt = sel_ms.contiguous(), sel_wgt.contiguous()
            else:
                idx = torch.tensor(active, device=dev)
                lms = sel_ms.index_select(2, idx).contiguous()
                lwt = sel_wgt.index_select(2, idx).contiguous()
            dirty = wgridder.ms2dirty(fuvw_d, sel_freq, lms, lwt, npixdirty, npixdirty, pixsize, pixsize, 0, 0,
                                      epsilon, do_wstacking, double_precision_accumulation=True,
                                      verbosity=verbosity, precision=precision)
            dirty = dirty.reshape(len(active), npixdirty, npixdirty).cpu().numpy()
            for k, p in enumerate(active):
                out[p] = dirty[k].T
        return out

    if mfs:
        dirty = grid(ms, wgt, freq, nonzero.any(axis=0))
        for pol in range(npol):
            if pol in dirty:
                im.pixels[0, pol] += dirty[pol]
            sumwt[0, pol] += numpy.sum(wsum[:, pol])
    else:
        for vchan in range(vnchan):
            ichan = vis_to_im[vchan]
            dirty = grid(ms[:, vchan : vchan + 1, :], wgt[:, vchan : vchan + 1, :], freq[vchan : vchan + 1],
                         nonzero[vchan])
            for pol in range(npol):
                if pol in dirty:
                    im.pixels[ichan, pol] += dirty[pol]
                sumwt[ichan, pol] += wsum[vchan, pol]

    if normalise:
        im = normalise_sumwt(im, sumwt)
    return im, sumwt
