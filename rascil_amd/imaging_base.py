"""Helpers around the imaging path: phase-centre shift and sum-of-weights normalisation.

Mirrors rascil/processing_components/imaging/base.py:67-162 (shift_vis_to_image,
normalise_sumwt 2-D case) and visibility/base.py:296-357, 1248-1264
(phaserotate_visibility, calculate_blockvisibility_phasor).
"""

import numpy

from .data_models import BlockVisibility, Direction, Image, copy_visibility, direction_to_lmn


def image_phasecentre(im: Image) -> Direction:
    """Direction of 0-relative pixel (nx//2, ny//2): the image's crval by construction."""
    return im.phasecentre


def calculate_blockvisibility_phasor(direction: Direction, vis: BlockVisibility):
    """exp(-2 pi i (u l + v m + w (n-1))) broadcast over pols (visibility/base.py:1248-1264)."""
    l, m, nm1 = direction_to_lmn(direction, vis.phasecentre)
    s = numpy.array([l, m, nm1])
    ph = numpy.exp(-2j * numpy.pi * numpy.einsum("tbfs,s->tbf", vis.uvw_lambda, s))
    return numpy.broadcast_to(ph[..., numpy.newaxis], vis.vis.shape)


def phaserotate_visibility(vis: BlockVisibility, newphasecentre: Direction, tangent=True,
                           inverse=False) -> BlockVisibility:
    """Phase rotate to a new phase centre (visibility/base.py:296-357).

    tangent=True adjusts only the phases (the imaging path's case). tangent=False also
    re-projects uvw (metres) onto the new phase centre, through the earth frame (uvw -> xyz at
    hour angle -ra of the old centre, xyz -> uvw at -ra of the new one), and moves the
    phase centre (:330-356)."""
    l, m, nm1 = direction_to_lmn(newphasecentre, vis.phasecentre)
    if numpy.abs(nm1) < 1e-15:
        return vis
    newvis = copy_visibility(vis)
    phasor = calculate_blockvisibility_phasor(newphasecentre, newvis)
    if inverse:
        newvis.vis = newvis.vis * phasor
    else:
        newvis.vis = newvis.vis * numpy.conj(phasor)
    if not tangent:
        from .simulation import uvw_to_xyz, xyz_to_uvw

        nrows, nbl, _ = vis.uvw.shape
        xyz = uvw_to_xyz(vis.uvw.reshape(nrows * nbl, 3), ha=-vis.phasecentre.ra, dec=vis.phasecentre.dec)
        uvw = xyz_to_uvw(xyz, ha=-newphasecentre.ra, dec=newphasecentre.dec)
        newvis.uvw = uvw.reshape(nrows, nbl, 3)
        newvis.phasecentre = newphasecentre
        newvis.attrs["phasecentre"] = newphasecentre
    return newvis


def shift_vis_to_image(vis: BlockVisibility, im: Image, tangent=True, inverse=False):
    """Shift vis to the image phase centre when they differ by > 1e-15 rad (imaging/base.py:67-106)."""
    ipc = image_phasecentre(im)
    if vis.phasecentre.separation(ipc) > 1e-15:
        vis = phaserotate_visibility(vis, ipc, tangent=tangent, inverse=inverse)
        vis.phasecentre = im.phasecentre
        vis.attrs["phasecentre"] = im.phasecentre
    return vis


def normalise_sumwt(im: Image, sumwt, min_weight=0.1, flat_sky=False) -> Image:
    """Divide each [chan, pol] plane by sumwt, or zero it if sumwt <= 0 (imaging/base.py:109-137)."""
    nchan, npol, _, _ = im.pixels.shape
    assert sumwt is not None
    sumwt = numpy.asarray(sumwt)
    if sumwt.shape != (nchan, npol):
        raise ValueError("sumwt is not a 2D array - cannot perform normalisation")
    for chan in range(nchan):
        for pol in range(npol):
            if sumwt[chan, pol] > 0.0:
                im.pixels[chan, pol, :, :] = im.pixels[chan, pol, :, :] / sumwt[chan, pol]
            else:
                im.pixels[chan, pol, :, :] = 0.0
    return im
