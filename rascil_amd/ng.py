"""invert_ng / predict_ng on the hipng GPU gridder.

Drop-in for rascil/processing_components/imaging/ng.py (predict_ng :33-133,
invert_ng :136-282): same signatures, keyword parameters (threads, epsilon,
do_wstacking, verbosity), conventions (u and w negated before the gridder, images in
RASCIL's [pol, y, x] pixel order, MFS when the image has one channel, per-channel otherwise
via the WCS frequency axis) and results. The ducc0.wgridder calls and the host passes
around them (copies, shift_vis_to_image, flag masking, convert_pol_frame, casts,
transposes, normalise_sumwt) become one gridder call per image channel that reads and
writes the visibility columns in place (hipng_geometry_invert_columns /
_predict_columns): all polarisations of a call in one batched pass.

Inputs: a BlockVisibility (numpy; uploaded per call) or a DeviceBlockVisibility (columns
resident in GPU memory; nothing crosses PCIe but the small metadata, and the gridder's
geometry for the same uvw, frequencies, image and accuracy is kept between calls, so a
predict after an invert of the same visibilities does not sort again). Image pixels may be
numpy or a GPU tensor.

Extra keyword: precision = "fp64" (default; the reference computes in FP64 with FP64
accumulation), "mixed" or "fp32".
"""

import collections
import logging

import numpy

from . import _lib, wgridder
from .data_models import DeviceBlockVisibility, direction_to_lmn, get_parameter, image_is_canonical
from .polarisation import pol_matrix

log = logging.getLogger("rascil-logger")

__all__ = ["predict_ng", "invert_ng", "clear_geometry_cache", "set_geometry_cache_size"]

_CACHE = collections.OrderedDict()
_CACHE_SIZE = [1]


def set_geometry_cache_size(n):
    """Number of gridder geometries kept for device-resident visibilities (0 disables)."""
    _CACHE_SIZE[0] = max(0, int(n))
    while len(_CACHE) > _CACHE_SIZE[0]:
        _CACHE.popitem(last=False)[1][1].close()


def clear_geometry_cache():
    """Release the cached geometries (and their device memory)."""
    while _CACHE:
        _CACHE.popitem(last=False)[1][1].close()


def _conv(ipf, opf):
    """Conversion matrix [opf.npol][ipf.npol] as interleaved (re, im) doubles: the layout the
    standalone passes hipng_vis_prologue / hipng_vis_epilogue take."""
    cm = numpy.ascontiguousarray(pol_matrix(ipf, opf), dtype=complex)
    return numpy.ascontiguousarray(cm.view(numpy.float64).ravel()), cm.shape


def _pixsize(im):
    return numpy.abs(numpy.radians(im.image_acc.wcs.cdelt[0]))


def _vis_to_im(model, freq):
    return numpy.round(model.image_acc.wcs.freq_to_pix(freq)).astype("int")


def _shift_lmn(vis_pc, im_pc):
    """Phase coefficients on the gridder's uvw for shift_vis_to_image (imaging/base.py:67-106
    with phaserotate_visibility visibility/base.py:296-357): the phase of RASCIL's uvw,
    u l + v m + w (n - 1), with u and w negated for the gridder, i.e. (-l, m, -(n - 1)).
    None when the centres coincide to 1e-15 rad or |n - 1| < 1e-15 (no rotation); `moved`
    says whether the visibility phase centre becomes the image's."""
    moved = vis_pc.separation(im_pc) > 1e-15
    if not moved:
        return None, False
    l, m, nm1 = direction_to_lmn(im_pc, vis_pc)
    if numpy.abs(nm1) < 1e-15:
        return None, True
    return (-l, m, -nm1), True


def _device_columns(torch, bvis, dev):
    """(vis, uvw [nrow, 3], imaging_weight, flags or None, cacheable) on the device."""
    if isinstance(bvis, DeviceBlockVisibility):
        return bvis.vis, bvis.uvw.reshape(-1, 3), bvis.imaging_weight, bvis.flags, True
    up = lambda a, dt: torch.from_numpy(numpy.ascontiguousarray(a)).to(device=dev, dtype=dt)  # noqa: E731
    flags = numpy.asarray(bvis.flags)
    return (up(bvis.vis, torch.complex128), up(bvis.uvw, torch.float64).reshape(-1, 3),
            up(bvis.imaging_weight, torch.float64), up(flags.astype(numpy.int8), torch.int8) if flags.any() else None,
            False)


def _geometry(torch, uvw, cacheable, freq, npix, pixsize, epsilon, do_wstacking, precision, nb):
    """The gridder geometry of the RASCIL call (u and w negated); cached for device columns,
    keyed on the uvw tensor (held by the cache, so its memory cannot be reused while cached)
    and its in-place version counter."""
    key = None
    if cacheable and _CACHE_SIZE[0] > 0:
        key = (uvw.data_ptr(), uvw._version, tuple(uvw.shape), str(uvw.device), freq.tobytes(), int(npix),
               float(pixsize), float(epsilon), bool(do_wstacking), precision, int(nb))
        hit = _CACHE.get(key)
        if hit is not None:
            _CACHE.move_to_end(key)
            return hit[1], False
    fuvw = uvw * torch.tensor([-1.0, 1.0, -1.0], dtype=torch.float64, device=uvw.device)
    g = wgridder.Geometry(fuvw, freq, npix, npix, pixsize, pixsize, epsilon, do_wstacking, precision, nbatch=nb,
                          device=uvw.device)
    del fuvw
    if key is None:
        return g, True
    _CACHE[key] = (uvw, g)
    while len(_CACHE) > _CACHE_SIZE[0]:
        _CACHE.popitem(last=False)[1][1].close()
    return g, False


def _image_dtype(torch, pix):
    if hasattr(pix, "dtype") and pix.dtype in (torch.float32, torch.float64):
        return pix.dtype
    return torch.float64


def predict_ng(bvis, model, **kwargs):
    """Predict visibilities from a model image (reference ng.py:33-133).

    One fused call per image channel: the degridder's unsort writes the visibility columns
    after the polarisation conversion to the visibility frame and the phase shift of
    shift_vis_to_image(inverse=True).

    :param bvis: BlockVisibility or DeviceBlockVisibility to be predicted (not modified)
    :param model: model image
    :return: new visibility set (same kind as bvis) with the predicted vis
    """
    assert image_is_canonical(model)
    if model is None:
        return bvis
    torch = wgridder._torch()
    epsilon = get_parameter(kwargs, "epsilon", 1e-12)
    do_wstacking = get_parameter(kwargs, "do_wstacking", True)
    verbosity = get_parameter(kwargs, "verbosity", 0)
    precision = get_parameter(kwargs, "precision", "fp64") or "fp64"

    on_device = isinstance(bvis, DeviceBlockVisibility)
    freq = numpy.ascontiguousarray(bvis.frequency, dtype="f8")
    nrows, nbaselines, vnchan, vnpol = bvis.vis.shape
    m_nchan, m_npol, ny, nx = model.pixels.shape
    assert m_npol == vnpol
    conv = pol_matrix(model.image_acc.polarisation_frame, bvis.blockvisibility_acc.polarisation_frame)
    lmn, moved = _shift_lmn(bvis.phasecentre, model.phasecentre)
    dev = bvis.vis.device if on_device else torch.device("cuda", torch.cuda.current_device())
    _, uvw, _, _, cacheable = _device_columns(torch, bvis, dev) if on_device else (
        None, torch.from_numpy(numpy.ascontiguousarray(bvis.uvw.reshape(-1, 3), "f8")).to(dev), None, None, False)
    vdt = bvis.vis.dtype if on_device else torch.complex128
    out = torch.empty((nrows * nbaselines, vnchan, vnpol), dtype=vdt, device=dev)
    pixsize = _pixsize(model)
    idt = _image_dtype(torch, model.pixels)
    pix = model.pixels if hasattr(model.pixels, "is_cuda") else torch.from_numpy(numpy.ascontiguousarray(model.pixels))
    groups = [(0, vnchan, 0)] if m_nchan == 1 else [(c, 1, int(i)) for c, i in enumerate(_vis_to_im(model, freq))]
    for chan0, nch, imchan in groups:
        g, transient = _geometry(torch, uvw, cacheable, freq[chan0:chan0 + nch], nx, pixsize, epsilon, do_wstacking,
                                 precision, m_npol)
        cols = wgridder.make_columns(vnchan, vnpol, conv, chan0=chan0, vis_dtype=vdt, lmn=lmn, image_order=1,
                                     image_dtype=idt)
        dirty = pix[imchan].to(device=dev, dtype=idt).contiguous()
        g.predict_columns(cols, dirty, out)
        if verbosity:
            print("hipng predict_ng plan:", g.plan)
        if transient:
            g.close()
    vis = out.reshape(nrows, nbaselines, vnchan, vnpol)
    if on_device:
        newbvis = bvis.with_vis(vis)
    else:
        newbvis = bvis.copy(zero=False)  # copy_visibility (visibility/base.py:105-118), vis replaced below
        newbvis.vis = vis.cpu().numpy()
    if moved:
        newbvis.phasecentre = model.phasecentre
        newbvis.attrs["phasecentre"] = model.phasecentre
    return newbvis


def invert_ng(bvis, model, dopsf: bool = False, normalise: bool = True, **kwargs):
    """Dirty image (or PSF) and sum of weights (reference ng.py:136-282).

    One fused call per image channel: the gridder's value gather reads the visibility
    columns with the flag mask, the phase shift to the image centre, the polarisation
    conversion and the imaging weights applied on the way, and the grid correction writes
    the image in RASCIL's pixel order already divided by sumwt (normalise); the input is not
    modified. Image polarisations whose visibilities are all zero (the reference skips them)
    come out zero as well.

    :param bvis: BlockVisibility or DeviceBlockVisibility to be inverted
    :param model: image template (not changed)
    :param dopsf: make the PSF instead of the dirty image
    :param normalise: normalise by the sum of weights
    :return: (image, sumwt [nchan, npol])
    """
    assert image_is_canonical(model)
    torch = wgridder._torch()
    lib = _lib.load()
    epsilon = get_parameter(kwargs, "epsilon", 1e-12)
    do_wstacking = get_parameter(kwargs, "do_wstacking", True)
    verbosity = get_parameter(kwargs, "verbosity", 0)
    precision = get_parameter(kwargs, "precision", "fp64") or "fp64"

    freq = numpy.ascontiguousarray(bvis.frequency, dtype="f8")
    nrows, nbaselines, vnchan, vnpol = bvis.vis.shape
    nchan, npol, ny, nx = model.pixels.shape
    conv = pol_matrix(bvis.blockvisibility_acc.polarisation_frame, model.image_acc.polarisation_frame)
    if conv.shape[0] != npol:
        raise ValueError("image polarisation frame does not match the image")
    if npol > vnpol:
        # the reference indexes the weights of the visibility frame by image polarisation
        raise IndexError(f"index {vnpol} is out of bounds for axis 2 with size {vnpol}")
    lmn, _ = _shift_lmn(bvis.phasecentre, model.phasecentre)
    on_device = isinstance(bvis, DeviceBlockVisibility)
    dev = bvis.vis.device if on_device else torch.device("cuda", torch.cuda.current_device())
    vis, uvw, iwt, flags, cacheable = _device_columns(torch, bvis, dev)
    n = nrows * nbaselines

    # sum of the flagged imaging weights per (channel, pol) of the visibility frame, FP64
    wsum_d = torch.empty((vnchan, vnpol), dtype=torch.float64, device=dev)
    _lib.check(lib.hipng_weight_sums(
        wgridder.ctypes.c_void_p(iwt.data_ptr()), wgridder.ctypes.c_void_p(flags.data_ptr()) if flags is not None
        else None, n, vnchan, vnpol, 0 if iwt.dtype == torch.float64 else 1, 1, wgridder.ctypes.c_void_p(
            wsum_d.data_ptr()), wgridder._stream(torch, dev)))
    wsum = wsum_d.cpu().numpy()
    vis_to_im = _vis_to_im(model, freq)
    mfs = nchan == 1 and vnchan > 1
    sumwt = numpy.zeros([nchan, npol])
    if mfs:
        sumwt[0, :] = wsum[:, :npol].sum(axis=0)
        groups = [(0, vnchan, 0)]
    else:
        for vchan in range(vnchan):
            sumwt[vis_to_im[vchan], :] += wsum[vchan, :npol]
        groups = [(c, 1, int(vis_to_im[c])) for c in range(vnchan)]

    im = model.copy(deep=True)
    pixsize = _pixsize(im)
    idt = _image_dtype(torch, im.pixels)
    on_device_image = hasattr(im.pixels, "is_cuda")
    if on_device_image:
        im.pixels.zero_()
    else:
        im.pixels[...] = 0.0
    for chan0, nch, ichan in groups:
        scale = [1.0 / sumwt[ichan, p] if sumwt[ichan, p] > 0.0 else 0.0 for p in range(npol)] if normalise else None
        g, transient = _geometry(torch, uvw, cacheable, freq[chan0:chan0 + nch], nx, pixsize, epsilon, do_wstacking,
                                 precision, npol)
        cols = wgridder.make_columns(vnchan, vnpol, conv, chan0=chan0, vis_dtype=vis.dtype, weight_dtype=iwt.dtype,
                                     flag_dtype=torch.int8, lmn=lmn, psf=dopsf, image_order=1, image_dtype=idt,
                                     scale=scale)
        if mfs and on_device_image:
            g.invert_columns(cols, vis, iwt, flags, out=im.pixels[0])
        else:
            dirty = g.invert_columns(cols, vis, iwt, flags)
            if on_device_image:
                im.pixels[ichan] += dirty
            else:
                im.pixels[ichan] += dirty.cpu().numpy()
        if verbosity:
            print("hipng invert_ng plan:", g.plan)
        if transient:
            g.close()
    return im, sumwt
