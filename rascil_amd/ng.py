"""invert_ng / predict_ng on the hipng GPU gridder.

Drop-in for rascil/processing_components/imaging/ng.py (predict_ng :33-133,
invert_ng :136-282): same signatures, keyword parameters (threads, epsilon,
do_wstacking, verbosity), conventions (u and w negated before the gridder, images
transposed to [x, y], MFS when the image has one channel, per-channel otherwise via the
WCS frequency axis) and results. The ducc0.wgridder calls are replaced by
rascil_amd.wgridder, which runs on the GPU and processes all polarisations of a call in
one batched pass instead of one call per polarisation.

Extra keyword: precision = "fp64" (default; the reference computes in FP64 with FP64
accumulation), "mixed" or "fp32".
"""

import logging

import numpy

from . import wgridder
from .data_models import BlockVisibility, Image, copy_visibility, get_parameter, image_is_canonical
from .imaging_base import normalise_sumwt, shift_vis_to_image
from .polarisation import convert_pol_frame

log = logging.getLogger("rascil-logger")

__all__ = ["predict_ng", "invert_ng"]


def _pixsize(im):
    return numpy.abs(numpy.radians(im.image_acc.wcs.cdelt[0]))


def _vis_to_im(model, freq):
    return numpy.round(model.image_acc.wcs.freq_to_pix(freq)).astype("int")


def _dtypes(precision):
    return ("c16", "f8") if precision == "fp64" else ("c8", "f4")


def predict_ng(bvis: BlockVisibility, model: Image, **kwargs) -> BlockVisibility:
    """Predict visibilities from a model image (reference ng.py:33-133).

    :param bvis: BlockVisibility to be predicted (not modified)
    :param model: model image
    :return: new BlockVisibility with the predicted vis
    """
    assert image_is_canonical(model)
    if model is None:
        return bvis
    epsilon = get_parameter(kwargs, "epsilon", 1e-12)
    do_wstacking = get_parameter(kwargs, "do_wstacking", True)
    verbosity = get_parameter(kwargs, "verbosity", 0)
    precision = get_parameter(kwargs, "precision", "fp64")
    cdt, rdt = _dtypes(precision)

    newbvis = copy_visibility(bvis, zero=True)
    freq = bvis.frequency
    nrows, nbaselines, vnchan, vnpol = bvis.vis.shape
    uvw = numpy.nan_to_num(newbvis.uvw.reshape([nrows * nbaselines, 3]))
    m_nchan, m_npol, ny, nx = model.pixels.shape
    assert m_npol == vnpol

    fuvw = uvw.copy()
    fuvw[:, 0] *= -1.0
    fuvw[:, 2] *= -1.0
    pixsize = _pixsize(model)
    vis_to_im = _vis_to_im(model, freq)

    vist = numpy.zeros([nrows * nbaselines, vnchan, vnpol], dtype=cdt)
    if m_nchan == 1:
        dirty = numpy.ascontiguousarray(model.pixels[0].transpose(0, 2, 1), dtype=rdt)
        vist[...] = wgridder.dirty2ms(fuvw, freq, dirty, None, pixsize, pixsize, 0, 0, epsilon,
                                      do_wstacking, precision=precision, verbosity=verbosity)
    else:
        for vchan in range(vnchan):
            imchan = vis_to_im[vchan]
            dirty = numpy.ascontiguousarray(model.pixels[imchan].transpose(0, 2, 1), dtype=rdt)
            vist[:, vchan : vchan + 1, :] = wgridder.dirty2ms(
                fuvw, freq[vchan : vchan + 1], dirty, None, pixsize, pixsize, 0, 0, epsilon,
                do_wstacking, precision=precision, verbosity=verbosity)

    vis = convert_pol_frame(vist.astype("c16"), model.image_acc.polarisation_frame,
                            bvis.blockvisibility_acc.polarisation_frame, polaxis=2)
    newbvis.vis = vis.reshape([nrows, nbaselines, vnchan, vnpol])
    return shift_vis_to_image(newbvis, model, tangent=True, inverse=True)


def invert_ng(bvis: BlockVisibility, model: Image, dopsf: bool = False, normalise: bool = True,
              **kwargs):
    """Dirty image (or PSF) and sum of weights (reference ng.py:136-282).

    :param bvis: BlockVisibility to be inverted
    :param model: image template (not changed)
    :param dopsf: make the PSF instead of the dirty image
    :param normalise: normalise by the sum of weights
    :return: (image, sumwt [nchan, npol])
    """
    assert image_is_canonical(model)
    im = model.copy(deep=True)
    epsilon = get_parameter(kwargs, "epsilon", 1e-12)
    do_wstacking = get_parameter(kwargs, "do_wstacking", True)
    verbosity = get_parameter(kwargs, "verbosity", 0)
    precision = get_parameter(kwargs, "precision", "fp64")
    cdt, rdt = _dtypes(precision)

    sbvis = copy_visibility(bvis)
    sbvis = shift_vis_to_image(sbvis, im, tangent=True, inverse=False)
    freq = sbvis.frequency
    nrows, nbaselines, vnchan, vnpol = sbvis.vis.shape

    ms = sbvis.blockvisibility_acc.flagged_vis.reshape([nrows * nbaselines, vnchan, vnpol])
    ms = convert_pol_frame(ms, bvis.blockvisibility_acc.polarisation_frame,
                           im.image_acc.polarisation_frame, polaxis=2).astype("c16")
    uvw = sbvis.uvw.reshape([nrows * nbaselines, 3])
    wgt = sbvis.blockvisibility_acc.flagged_imaging_weight.astype("f8")
    wgt = wgt.reshape([nrows * nbaselines, vnchan, vnpol])

    npixdirty = im.pixels.shape[-1]
    pixsize = _pixsize(im)
    fuvw = uvw.copy()
    fuvw[:, 0] *= -1.0
    fuvw[:, 2] *= -1.0

    nchan, npol, ny, nx = im.pixels.shape
    im.pixels[...] = 0.0
    sumwt = numpy.zeros([nchan, npol])
    vis_to_im = _vis_to_im(model, freq)
    mfs = nchan == 1 and vnchan > 1
    if dopsf:
        ms[...] = 0.0
        ms[..., 0] = 1.0

    def grid(sel_ms, sel_wgt, sel_freq):
        # polarisations with data are gridded together; all-zero ones are skipped as in the
        # reference (their weights still enter sumwt)
        active = [p for p in range(npol) if numpy.max(numpy.abs(sel_ms[..., p])) > 0.0]
        out = {}
        if active:
            lms = numpy.ascontiguousarray(sel_ms[..., active], dtype=cdt)
            lwt = numpy.ascontiguousarray(sel_wgt[..., active], dtype=rdt)
            dirty = wgridder.ms2dirty(fuvw, sel_freq, lms, lwt, npixdirty, npixdirty, pixsize,
                                      pixsize, 0, 0, epsilon, do_wstacking,
                                      double_precision_accumulation=True, verbosity=verbosity,
                                      precision=precision)
            for k, p in enumerate(active):
                out[p] = dirty[k].T
        return out

    if mfs:
        dirty = grid(ms, wgt, freq)
        for pol in range(npol):
            if pol in dirty:
                im.pixels[0, pol] += dirty[pol]
            sumwt[0, pol] += numpy.sum(wgt[:, :, pol])
    else:
        for vchan in range(vnchan):
            ichan = vis_to_im[vchan]
            dirty = grid(ms[:, vchan : vchan + 1, :], wgt[:, vchan : vchan + 1, :],
                         freq[vchan : vchan + 1])
            for pol in range(npol):
                if pol in dirty:
                    im.pixels[ichan, pol] += dirty[pol]
                sumwt[ichan, pol] += numpy.sum(wgt[:, vchan, pol])

    if normalise:
        im = normalise_sumwt(im, sumwt)
    return im, sumwt
