"""Polarisation conversion on the pol axis (rascil/data_models/polarisation.py:164-419).

convert_pol_frame applies the same conversion matrices as the reference (TMS eq. 4.58 /
4.59 and their inverses, :220-360) as one complex matrix on `polaxis`. The *->stokesI
reductions of the reference act on the last axis only; here they act on `polaxis`,
which coincides with the reference for the visibility layout [row, chan, pol].
"""

import numpy

from .data_models import PolarisationFrame

_S2L4 = numpy.array([[1, 1, 0, 0], [0, 0, 1, 1j], [0, 0, 1, -1j], [1, -1, 0, 0]])
_S2C4 = numpy.array([[1, 0, 0, 1], [0, -1j, 1, 0], [0, -1j, -1, 0], [1, 0, 0, -1]])
_S2X2 = numpy.array([[1, 1], [1, -1]])
_L2S4 = numpy.array(
    [[0.5, 0, 0, 0.5], [0.5, 0, 0, -0.5], [0, 0.5, 0.5, 0], [0, -0.5j, 0.5j, 0]]
)
_C2S4 = numpy.array(
    [[0.5, 0, 0, 0.5], [0, 0.5j, 0.5j, 0], [0, 0.5, -0.5, 0], [0.5, 0, 0, -0.5]]
)
_X2S2 = numpy.array([[0.5, 0.5], [0.5, -0.5]])


def pol_matrix(ipf: PolarisationFrame, opf: PolarisationFrame):
    """Complex matrix [opf.npol, ipf.npol] converting ipf to opf; ValueError if unsupported."""
    i, o = ipf.type, opf.type
    if i == o:
        return numpy.eye(ipf.npol, dtype=complex)
    table = {
        ("linear", "stokesIQUV"): _L2S4,
        ("linear", "stokesI"): numpy.array([[0.5, 0, 0, 0.5]]),
        ("linearnp", "stokesIQ"): _X2S2,
        ("linearnp", "stokesI"): numpy.array([[0.5, 0.5]]),
        ("circular", "stokesIQUV"): _C2S4,
        ("circular", "stokesI"): numpy.array([[0.5, 0, 0, 0.5]]),
        ("circularnp", "stokesIV"): _X2S2,
        ("circularnp", "stokesI"): numpy.array([[0.5, 0.5]]),
        ("stokesIQUV", "linear"): _S2L4,
        ("stokesIQUV", "circular"): _S2C4,
        ("stokesIQUV", "stokesI"): numpy.array([[1, 0, 0, 0]]),
        ("stokesIQ", "linearnp"): _S2X2,
        ("stokesIV", "circularnp"): _S2X2,
        ("stokesI", "stokesIQUV"): numpy.array([[1], [0], [0], [0]]),
    }
    if (i, o) not in table:
        raise ValueError("Unknown polarisation conversion: {} to {}".format(ipf, opf))
    return numpy.asarray(table[(i, o)], dtype=complex)


def convert_pol_frame(polvec, ipf: PolarisationFrame, opf: PolarisationFrame, polaxis=1):
    """Convert `polvec` from frame ipf to opf along `polaxis`."""
    if ipf == opf:
        return polvec
    cm = pol_matrix(ipf, opf)
    polvec = numpy.asarray(polvec)
    if polaxis < 0:
        polaxis += polvec.ndim
    moved = numpy.moveaxis(polvec, polaxis, -1)
    out = moved @ cm.T
    return numpy.moveaxis(out, -1, polaxis)
