"""ORACLE — test infrastructure only (never imported by rascil_amd).

FP64 brute-force restatements of the reference path:

* explicit_ms2dirty / explicit_dirty2ms: the exact sums ducc0.wgridder.ms2dirty /
  dirty2ms approximate (ducc0 0.24, pinned at requirements.txt:83; not vendored, not
  installed, so restated from its published definition), as RASCIL calls them from
  rascil/processing_components/imaging/ng.py:91-120, 232-275:
      dirty(x, y) = sum_{r,c} wgt Re[ms exp(+2 pi i f/c (u x + v y - w (n-1)))] / n
      ms(r, c)    = wgt sum_{x,y} dirty(x, y) / n exp(-2 pi i f/c (u x + v y - w (n-1)))
  with x = (ix - nx/2) px, y = (iy - ny/2) py, n - 1 = -r^2/(sqrt(1-r^2)+1) and, without
  w-stacking, n = 1 and no w term.
* dft_cpu_looped: rascil/processing_components/imaging/dft.py:261-281.
* pol_matrix / convert_pol_frame: rascil/data_models/polarisation.py:220-419.
* phasor: rascil/processing_components/visibility/base.py:1248-1264.

Pinning (tests/test_oracle.py): the DFT known answers of
tests/processing_components/test_dft_skycomponent_visibility_kernels.py:83-85, 124-126;
the predict unit-pixel invariants of tests/processing_components/test_imaging.py:274-293;
the PSF-peak invariant (:601-626); the adjoint identity between the two explicit sums;
closed forms for single visibilities; and hermitian symmetry.
"""

import numpy

C_M_S = 299792458.0  # rascil/phyconst.py:6


def _lm_grid(npix_x, npix_y, px, py, do_wstacking):
    x = (numpy.arange(npix_x) - npix_x // 2) * px
    y = (numpy.arange(npix_y) - npix_y // 2) * py
    xx, yy = numpy.meshgrid(x, y, indexing="ij")
    if do_wstacking:
        r2 = xx ** 2 + yy ** 2
        nm1 = -r2 / (numpy.sqrt(1.0 - r2) + 1.0)
        n = nm1 + 1.0
    else:
        nm1 = numpy.zeros_like(xx)
        n = numpy.ones_like(xx)
    return xx, yy, nm1, n


def _flat_lambda(uvw, freq):
    f = numpy.asarray(freq, "f8") / C_M_S
    u = (uvw[:, 0:1] * f[None, :]).ravel()
    v = (uvw[:, 1:2] * f[None, :]).ravel()
    w = (uvw[:, 2:3] * f[None, :]).ravel()
    return u, v, w


def explicit_ms2dirty(uvw, freq, ms, wgt, npix_x, npix_y, pixsize_x, pixsize_y, do_wstacking):
    """Exact adjoint sum; ms [nrow, nchan] complex, wgt [nrow, nchan] or None -> [nx, ny]."""
    xx, yy, nm1, n = _lm_grid(npix_x, npix_y, pixsize_x, pixsize_y, do_wstacking)
    u, v, w = _flat_lambda(numpy.asarray(uvw, "f8"), freq)
    a = numpy.asarray(ms, "c16").ravel()
    if wgt is not None:
        a = a * numpy.asarray(wgt, "f8").ravel()
    res = numpy.zeros((npix_x, npix_y))
    chunk = max(1, (1 << 21) // (npix_x * npix_y))
    for s in range(0, a.size, chunk):
        sl = slice(s, s + chunk)
        ph = (u[sl, None, None] * xx[None] + v[sl, None, None] * yy[None]
              - w[sl, None, None] * nm1[None])
        res += numpy.einsum("k,kxy->xy", a[sl], numpy.exp(2j * numpy.pi * ph)).real
    return res / n


def explicit_ms2dirty_at(uvw, freq, ms, wgt, ix, iy, npix_x, npix_y, pixsize_x, pixsize_y, do_wstacking):
    """explicit_ms2dirty at the pixels (ix[k], iy[k]) of an npix_x x npix_y image only (for
    images too large to sum whole)."""
    ix = numpy.asarray(ix, "i8")
    iy = numpy.asarray(iy, "i8")
    x = (ix - npix_x // 2) * pixsize_x
    y = (iy - npix_y // 2) * pixsize_y
    if do_wstacking:
        r2 = x ** 2 + y ** 2
        nm1 = -r2 / (numpy.sqrt(1.0 - r2) + 1.0)
    else:
        nm1 = numpy.zeros_like(x)
    u, v, w = _flat_lambda(numpy.asarray(uvw, "f8"), freq)
    a = numpy.asarray(ms, "c16").ravel()
    if wgt is not None:
        a = a * numpy.asarray(wgt, "f8").ravel()
    ph = u[:, None] * x[None] + v[:, None] * y[None] - w[:, None] * nm1[None]
    return (a @ numpy.exp(2j * numpy.pi * ph)).real / (nm1 + 1.0)


def explicit_dirty2ms(uvw, freq, dirty, wgt, pixsize_x, pixsize_y, do_wstacking):
    """Exact forward sum; dirty [nx, ny] -> ms [nrow, nchan] complex128."""
    dirty = numpy.asarray(dirty, "f8")
    npix_x, npix_y = dirty.shape
    xx, yy, nm1, n = _lm_grid(npix_x, npix_y, pixsize_x, pixsize_y, do_wstacking)
    u, v, w = _flat_lambda(numpy.asarray(uvw, "f8"), freq)
    d = (dirty / n).ravel()
    out = numpy.zeros(u.size, "c16")
    chunk = max(1, (1 << 21) // (npix_x * npix_y))
    for s in range(0, u.size, chunk):
        sl = slice(s, s + chunk)
        ph = (u[sl, None, None] * xx[None] + v[sl, None, None] * yy[None]
              - w[sl, None, None] * nm1[None])
        out[sl] = numpy.exp(-2j * numpy.pi * ph).reshape(ph.shape[0], -1) @ d
    out = out.reshape(uvw.shape[0], len(numpy.atleast_1d(freq)))
    if wgt is not None:
        out = out * numpy.asarray(wgt, "f8").reshape(out.shape)
    return out


def explicit_dirty2ms_sparse(uvw, freq, ix, iy, values, npix_x, npix_y, pixsize_x, pixsize_y, wgt,
                            do_wstacking):
    """explicit_dirty2ms of an npix_x x npix_y image that is zero except dirty[ix[k], iy[k]] =
    values[k] (for images too large to sum whole)."""
    ix = numpy.asarray(ix, "i8")
    iy = numpy.asarray(iy, "i8")
    x = (ix - npix_x // 2) * pixsize_x
    y = (iy - npix_y // 2) * pixsize_y
    if do_wstacking:
        r2 = x ** 2 + y ** 2
        nm1 = -r2 / (numpy.sqrt(1.0 - r2) + 1.0)
    else:
        nm1 = numpy.zeros_like(x)
    u, v, w = _flat_lambda(numpy.asarray(uvw, "f8"), freq)
    ph = u[:, None] * x[None] + v[:, None] * y[None] - w[:, None] * nm1[None]
    out = numpy.exp(-2j * numpy.pi * ph) @ (numpy.asarray(values, "f8") / (nm1 + 1.0))
    out = out.reshape(uvw.shape[0], len(numpy.atleast_1d(freq)))
    if wgt is not None:
        out = out * numpy.asarray(wgt, "f8").reshape(out.shape)
    return out


def dft_cpu_looped(direction_cosines, uvw_lambda, vfluxes):
    """vis[t,b,c,p] = sum_k F[k,c,p] exp(-2 pi i uvw_lambda . (l, m, n-1)_k) (dft.py:261-281)."""
    ncomp = direction_cosines.shape[0]
    ntimes, nbaselines, nchan, _ = uvw_lambda.shape
    npol = vfluxes.shape[-1]
    vis = numpy.zeros([ntimes, nbaselines, nchan, npol], dtype="complex")
    for icomp in range(ncomp):
        phasor = numpy.exp(-2j * numpy.pi * numpy.sum(uvw_lambda * direction_cosines[icomp, :], axis=-1))
        for pol in range(npol):
            vis[..., pol] += vfluxes[icomp, :, pol] * phasor
    return vis


# rascil/data_models/polarisation.py:240-264 (stokes<->linear), :300-324 (stokes<->circular)
POL_MATRICES = {
    ("stokesIQUV", "linear"): numpy.array([[1, 1, 0, 0], [0, 0, 1, 1j], [0, 0, 1, -1j], [1, -1, 0, 0]]),
    ("linear", "stokesIQUV"): numpy.array([[0.5, 0, 0, 0.5], [0.5, 0, 0, -0.5], [0, 0.5, 0.5, 0],
                                           [0, -0.5j, 0.5j, 0]]),
    ("stokesIQUV", "circular"): numpy.array([[1, 0, 0, 1], [0, -1j, 1, 0], [0, -1j, -1, 0], [1, 0, 0, -1]]),
    ("circular", "stokesIQUV"): numpy.array([[0.5, 0, 0, 0.5], [0, 0.5j, 0.5j, 0], [0, 0.5, -0.5, 0],
                                             [0.5, 0, 0, -0.5]]),
    ("linear", "stokesI"): numpy.array([[0.5, 0, 0, 0.5]]),
    ("circular", "stokesI"): numpy.array([[0.5, 0, 0, 0.5]]),
    ("stokesIQUV", "stokesI"): numpy.array([[1, 0, 0, 0]]),
    ("stokesI", "stokesIQUV"): numpy.array([[1], [0], [0], [0]]),
    # two-polarisation frames (polarisation.py:229-231, 250-256, 275-276, 290-292, 310-316,
    # 336-337 and the dispatch :379-409): stokesIQ <-> linearnp (XX, YY) and stokesIV <->
    # circularnp (RR, LL), and their reductions to Stokes I
    ("stokesIQ", "linearnp"): numpy.array([[1, 1], [1, -1]]),
    ("linearnp", "stokesIQ"): numpy.array([[0.5, 0.5], [0.5, -0.5]]),
    ("stokesIV", "circularnp"): numpy.array([[1, 1], [1, -1]]),
    ("circularnp", "stokesIV"): numpy.array([[0.5, 0.5], [0.5, -0.5]]),
    ("linearnp", "stokesI"): numpy.array([[0.5, 0.5]]),
    ("circularnp", "stokesI"): numpy.array([[0.5, 0.5]]),
}


def convert_pol(vec, ipf, opf, polaxis=-1):
    """Apply the literal conversion matrix on `polaxis` (ipf/opf frame names)."""
    if ipf == opf:
        return vec
    cm = POL_MATRICES[(ipf, opf)].astype(complex)
    return numpy.moveaxis(numpy.tensordot(cm, vec, axes=(1, polaxis)), 0, polaxis)


def phasor(uvw_lambda, l, m, nm1):
    """exp(-2 pi i (u l + v m + w (n-1))) over [t, b, c] (visibility/base.py:1248-1264)."""
    return numpy.exp(-2j * numpy.pi * (uvw_lambda[..., 0] * l + uvw_lambda[..., 1] * m
                                       + uvw_lambda[..., 2] * nm1))
