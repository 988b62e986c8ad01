"""ORACLE — test infrastructure only (never imported by rascil_amd).

numpy restatement of RASCIL's imaging weights: grid_blockvisibility_weight_to_griddata
(rascil/processing_components/griddata/gridding.py:260-308) with the nearest-cell mapping
of spatial_mapping (:154-176) on the grid WCS of create_griddata_from_image
(griddata/operations.py:71-107), and griddata_blockvisibility_reweight (:336-424).
"""

import numpy

C_M_S = 299792458.0


def cells(uvw, freq, nx, ny, cdelt0_deg, cdelt1_deg):
    d2r = numpy.pi / 180.0
    du = 1.0 / (nx * cdelt0_deg * d2r)
    dv = 1.0 / (ny * cdelt1_deg * d2r)
    fc = numpy.asarray(freq) / C_M_S
    u = uvw[:, 0:1] * fc[None, :]
    v = uvw[:, 1:2] * fc[None, :]
    pu = numpy.round(u / du + (nx // 2 + 1) - 1.0).astype(int)
    pv = numpy.round(v / dv + (ny // 2 + 1) - 1.0).astype(int)
    assert pu.min() >= 0 and pu.max() < nx and pv.min() >= 0 and pv.max() < ny
    return pu, pv


def weight(uvw, freq, imwt, vis_to_im, nimchan, nx, ny, cdelt0_deg, cdelt1_deg, weighting, robustness=0.0):
    """imwt [nrow, nchan, npol] flagged imaging weights -> (new weights, grid)."""
    nrow, nchan, npol = imwt.shape
    pu, pv = cells(uvw, freq, nx, ny, cdelt0_deg, cdelt1_deg)
    grid = numpy.zeros((nimchan, npol, ny, nx))
    for ch in range(nchan):
        for p in range(npol):
            numpy.add.at(grid[vis_to_im[ch], p], (pv[:, ch], pu[:, ch]), imwt[:, ch, p])
    out = imwt.copy()
    sumlocwt = numpy.sum(grid ** 2)
    sumwt = numpy.sum(imwt)
    f2 = (5.0 * numpy.power(10.0, -robustness)) ** 2 * sumwt / sumlocwt
    for ch in range(nchan):
        for p in range(npol):
            g = grid[vis_to_im[ch], p, pv[:, ch], pu[:, ch]]
            w = out[:, ch, p]
            pos = g > 0
            if weighting == "uniform":
                w[pos] *= w[pos] / g[pos]
            else:
                w[pos] *= (1 + f2 * w[pos]) / (1 + f2 * g[pos])
            w[~pos] = 0.0
    return out, grid
