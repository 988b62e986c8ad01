"""Lightweight data models carrying the fields RASCIL's imaging path consumes.

RASCIL's BlockVisibility and Image are xarray Datasets with astropy WCS
(rascil/data_models/memory_data_models.py:511-718, :1214-1529). This package keeps the
same field names, shapes, dtypes and accessors that invert_ng / predict_ng touch, as
plain numpy-backed classes:

* BlockVisibility: vis [time, baseline, chan, pol] complex128, uvw [time, baseline, 3]
  metres, weight / imaging_weight [time, baseline, chan, pol] float64, flags int,
  frequency [chan] Hz, channel_bandwidth, polarisation_frame, phasecentre.
* Image: pixels [chan, pol, ny, nx] float64 with a SIN-projection WCS in the RASCIL
  convention (cdelt[0] = -cell, crpix = N//2 + 1, crval = phase centre; frequency axis 4).

Directions are (ra, dec) in radians (Direction), replacing astropy SkyCoord.
"""

import copy
import math
from dataclasses import dataclass

import numpy

C_M_S = 299792458.0  # rascil/phyconst.py:6


class PolarisationFrame:
    """Polarisation frames post correlation (rascil/data_models/polarisation.py:78-150)."""

    polarisation_frames = {
        "circular": {"RR": 0, "RL": 1, "LR": 2, "LL": 3},
        "circularnp": {"RR": 0, "LL": 1},
        "linear": {"XX": 0, "XY": 1, "YX": 2, "YY": 3},
        "linearnp": {"XX": 0, "YY": 1},
        "stokesIQUV": {"I": 0, "Q": 1, "U": 2, "V": 3},
        "stokesIV": {"I": 0, "V": 1},
        "stokesIQ": {"I": 0, "Q": 1},
        "stokesI": {"I": 0},
    }

    def __init__(self, name):
        if name not in self.polarisation_frames:
            raise ValueError("Unknown polarisation frame %s" % str(name))
        self.type = name
        self.translations = self.polarisation_frames[name]

    def __eq__(self, a):
        if a is None:
            return False
        return self.type == a.type

    def __hash__(self):
        return hash(self.type)

    def __str__(self):
        return self.type

    def __repr__(self):
        return f"PolarisationFrame({self.type!r})"

    @property
    def npol(self):
        return len(self.translations)

    @property
    def names(self):
        return list(self.translations.keys())


@dataclass(frozen=True)
class Direction:
    """A sky direction (ra, dec) in radians; replaces astropy SkyCoord on this path."""

    ra: float
    dec: float

    def separation(self, other):
        """Angular separation in radians (haversine form)."""
        d_ra = other.ra - self.ra
        s = math.sin((other.dec - self.dec) / 2) ** 2 + math.cos(self.dec) * math.cos(
            other.dec
        ) * math.sin(d_ra / 2) ** 2
        return 2 * math.asin(min(1.0, math.sqrt(s)))


def direction_to_lmn(pos: Direction, phasecentre: Direction):
    """(l, m, n-1) of `pos` relative to `phasecentre`, l increasing east-wards.

    Same quantities as skycoord_to_lmn (rascil/processing_components/util/
    coordinate_support.py:446-465): astropy's sky-offset-frame cartesian (y, z, x - 1).
    """
    dra = pos.ra - phasecentre.ra
    l = math.cos(pos.dec) * math.sin(dra)
    m = math.sin(pos.dec) * math.cos(phasecentre.dec) - math.cos(pos.dec) * math.sin(
        phasecentre.dec
    ) * math.cos(dra)
    n = math.sin(pos.dec) * math.sin(phasecentre.dec) + math.cos(pos.dec) * math.cos(
        phasecentre.dec
    ) * math.cos(dra)
    return l, m, n - 1.0


def lmn_to_direction(l, m, phasecentre: Direction):
    """Inverse of direction_to_lmn for a point on the unit sphere (n from l, m)."""
    n = math.sqrt(1.0 - l * l - m * m)
    sd, cd = math.sin(phasecentre.dec), math.cos(phasecentre.dec)
    dec = math.asin(m * cd + n * sd)
    ra = phasecentre.ra + math.atan2(l, n * cd - m * sd)
    return Direction(ra, dec)


class _Field:
    """`obj["name"].data` accessor shim (xarray-style)."""

    def __init__(self, owner, name):
        self._owner = owner
        self._name = name

    @property
    def data(self):
        return getattr(self._owner, self._name)

    @data.setter
    def data(self, value):
        setattr(self._owner, self._name, value)

    @property
    def shape(self):
        return getattr(self._owner, self._name).shape


class _BVisAccessor:
    """bvis.blockvisibility_acc (memory_data_models.py:1448-1529)."""

    def __init__(self, bvis):
        self._b = bvis

    @property
    def polarisation_frame(self):
        return self._b.polarisation_frame

    @property
    def flagged_vis(self):
        return self._b.vis * (1 - self._b.flags)

    @property
    def flagged_weight(self):
        return self._b.weight * (1 - self._b.flags)

    @property
    def flagged_imaging_weight(self):
        return self._b.imaging_weight * (1 - self._b.flags)

    @property
    def nvis(self):
        return self._b.vis.size

    @property
    def uvw_lambda(self):
        return self._b.uvw_lambda


class BlockVisibility:
    """Visibility block [time, baseline, chan, pol] with the RASCIL field names."""

    def __init__(
        self,
        vis,
        uvw,
        frequency,
        polarisation_frame,
        phasecentre,
        weight=None,
        imaging_weight=None,
        flags=None,
        channel_bandwidth=None,
        time=None,
        baselines=None,
    ):
        self.vis = numpy.asarray(vis)
        nt, nb, nc, npol = self.vis.shape
        self.uvw = numpy.asarray(uvw, dtype="f8").reshape(nt, nb, 3)
        self.frequency = numpy.asarray(frequency, dtype="f8").reshape(nc)
        self.polarisation_frame = polarisation_frame
        assert polarisation_frame.npol == npol, "pol frame does not match vis shape"
        self.phasecentre = phasecentre
        shape = self.vis.shape
        self.weight = numpy.ones(shape) if weight is None else numpy.asarray(weight, "f8")
        self.imaging_weight = (
            self.weight.copy() if imaging_weight is None else numpy.asarray(imaging_weight, "f8")
        )
        self.flags = numpy.zeros(shape, "int") if flags is None else numpy.asarray(flags, "int")
        self.channel_bandwidth = (
            numpy.ones(nc) * 1e6 if channel_bandwidth is None else numpy.asarray(channel_bandwidth)
        )
        self.time = numpy.arange(nt, dtype="f8") if time is None else numpy.asarray(time)
        self.baselines = baselines
        self.attrs = {"phasecentre": phasecentre}

    def __getitem__(self, name):
        return _Field(self, name)

    @property
    def blockvisibility_acc(self):
        return _BVisAccessor(self)

    @property
    def uvw_lambda(self):
        """uvw * f / c, [time, baseline, chan, 3] (memory_data_models.py:1315-1320)."""
        k = self.frequency / C_M_S
        return numpy.einsum("tbs,k->tbks", self.uvw, k)

    @property
    def nvis(self):
        return self.vis.size

    def copy(self, zero=False):
        new = copy.deepcopy(self)
        if zero:
            new.vis[...] = 0.0
        return new


class DeviceBlockVisibility:
    """A BlockVisibility whose columns live in GPU memory as torch tensors.

    Same fields, shapes and accessors as BlockVisibility (memory_data_models.py:1214-1529):
    vis [time, baseline, chan, pol] (complex128, or complex64 to halve the memory),
    uvw [time, baseline, 3] float64 metres, weight / imaging_weight [time, baseline, chan, pol]
    (float64 or float32), flags int8 (or None when nothing is flagged), frequency and the
    other small metadata on the host. invert_ng / predict_ng read and write these columns in
    place on the device: no host copies, and the geometry (sort, records, plans) of a
    (uvw, frequency, image, accuracy) combination is kept between calls. copy() shares
    every column but vis (columns are treated as immutable; replace them, do not modify
    them in place)."""

    def __init__(self, vis, uvw, frequency, polarisation_frame, phasecentre, weight=None, imaging_weight=None,
                 flags=None, channel_bandwidth=None, time=None, baselines=None):
        import torch

        if not (hasattr(vis, "is_cuda") and vis.is_cuda):
            raise ValueError("DeviceBlockVisibility needs GPU tensors (use from_host for numpy data)")
        nt, nb, nc, npol = vis.shape
        if vis.dtype not in (torch.complex128, torch.complex64) or not vis.is_contiguous():
            raise ValueError("vis must be a contiguous complex128 or complex64 tensor")
        assert polarisation_frame.npol == npol, "pol frame does not match vis shape"
        self.vis = vis
        self.uvw = uvw.reshape(nt, nb, 3).to(device=vis.device, dtype=torch.float64).contiguous()
        self.frequency = numpy.asarray(frequency, dtype="f8").reshape(nc)
        self.polarisation_frame = polarisation_frame
        self.phasecentre = phasecentre
        rdt = torch.float64 if vis.dtype == torch.complex128 else torch.float32
        self.weight = (torch.ones(vis.shape, dtype=rdt, device=vis.device) if weight is None
                       else weight.to(vis.device).contiguous())
        self.imaging_weight = self.weight if imaging_weight is None else imaging_weight.to(vis.device).contiguous()
        for name in ("weight", "imaging_weight"):
            t = getattr(self, name)
            if tuple(t.shape) != tuple(vis.shape) or t.dtype not in (torch.float64, torch.float32):
                raise ValueError(f"{name} must be float64 or float32 of the vis shape")
        self.flags = None if flags is None else flags.to(device=vis.device, dtype=torch.int8).contiguous()
        if self.flags is not None and tuple(self.flags.shape) != tuple(vis.shape):
            raise ValueError("flags must have the vis shape")
        self.channel_bandwidth = (numpy.ones(nc) * 1e6 if channel_bandwidth is None
                                  else numpy.asarray(channel_bandwidth))
        self.time = numpy.arange(nt, dtype="f8") if time is None else numpy.asarray(time)
        self.baselines = baselines
        self.attrs = {"phasecentre": phasecentre}

    @classmethod
    def from_host(cls, bvis: BlockVisibility, device=None, low_precision=False):
        """Upload a BlockVisibility; low_precision stores vis complex64 and weights float32."""
        import torch

        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        cdt = torch.complex64 if low_precision else torch.complex128
        rdt = torch.float32 if low_precision else torch.float64
        up = lambda a, dt: torch.from_numpy(numpy.ascontiguousarray(a)).to(device=device, dtype=dt)  # noqa: E731
        flags = numpy.asarray(bvis.flags)
        return cls(up(bvis.vis, cdt), up(bvis.uvw, torch.float64), bvis.frequency, bvis.polarisation_frame,
                   bvis.phasecentre, weight=up(bvis.weight, rdt), imaging_weight=up(bvis.imaging_weight, rdt),
                   flags=up(flags.astype(numpy.int8), torch.int8) if flags.any() else None,
                   channel_bandwidth=bvis.channel_bandwidth, time=bvis.time, baselines=bvis.baselines)

    def to_host(self) -> BlockVisibility:
        """The columns copied back into a (numpy) BlockVisibility."""
        dn = lambda t: t.detach().cpu().numpy()  # noqa: E731
        b = BlockVisibility(dn(self.vis).astype("c16"), dn(self.uvw), self.frequency, self.polarisation_frame,
                            self.phasecentre, weight=dn(self.weight).astype("f8"),
                            imaging_weight=dn(self.imaging_weight).astype("f8"),
                            flags=None if self.flags is None else dn(self.flags).astype("int"),
                            channel_bandwidth=self.channel_bandwidth, time=self.time, baselines=self.baselines)
        b.attrs = dict(self.attrs)
        return b

    def __getitem__(self, name):
        return _Field(self, name)

    @property
    def blockvisibility_acc(self):
        return _DeviceBVisAccessor(self)

    @property
    def nvis(self):
        return self.vis.numel()

    @property
    def uvw_lambda(self):
        """uvw * f / c, [time, baseline, chan, 3] float64 on the device."""
        import torch

        k = torch.tensor(self.frequency / C_M_S, dtype=torch.float64, device=self.vis.device)
        return self.uvw[..., None, :] * k[None, None, :, None]

    def with_vis(self, vis):
        """A new visibility set with the given vis tensor, sharing every other column."""
        if tuple(vis.shape) != tuple(self.vis.shape) or vis.device != self.vis.device:
            raise ValueError("vis must have the shape and device of the visibility set")
        new = copy.copy(self)
        new.attrs = dict(self.attrs)
        new.vis = vis.contiguous()
        return new

    def copy(self, zero=False):
        """A new visibility set sharing every column but vis (a copy, or zeros)."""
        import torch

        new = copy.copy(self)
        new.attrs = dict(self.attrs)
        new.vis = torch.zeros_like(self.vis) if zero else self.vis.clone()
        return new


class _DeviceBVisAccessor(_BVisAccessor):
    """blockvisibility_acc of a DeviceBlockVisibility (torch tensors)."""

    def _keep(self):
        f = self._b.flags
        return 1 if f is None else (1 - f.to(self._b.weight.dtype))

    @property
    def flagged_vis(self):
        return self._b.vis * self._keep()

    @property
    def flagged_weight(self):
        return self._b.weight * self._keep()

    @property
    def flagged_imaging_weight(self):
        return self._b.imaging_weight * self._keep()


def copy_visibility(vis, zero=False):
    """Deep copy, optionally zeroing vis (rascil/processing_components/visibility/base.py:105-118);
    a DeviceBlockVisibility shares its other columns (see DeviceBlockVisibility.copy)."""
    return vis.copy(zero=zero)


class _WCS:
    """The four-axis SIN/STOKES/FREQ WCS numbers RASCIL's path reads (cdelt, crpix, crval)."""

    def __init__(self, cdelt, crpix, crval, ctype):
        self.cdelt = numpy.array(cdelt, dtype="f8")
        self.crpix = numpy.array(crpix, dtype="f8")
        self.crval = numpy.array(crval, dtype="f8")
        self.ctype = list(ctype)

    def freq_to_pix(self, freq):
        """0-relative pixel on the frequency axis (wcs.sub([4]).wcs_world2pix(freq, 0))."""
        return (numpy.asarray(freq) - self.crval[3]) / self.cdelt[3] + self.crpix[3] - 1.0


class _ImageAccessor:
    def __init__(self, im):
        self._im = im

    @property
    def polarisation_frame(self):
        return self._im.polarisation_frame

    @property
    def shape(self):
        return self._im.pixels.shape

    @property
    def phasecentre(self):
        return self._im.phasecentre

    @property
    def wcs(self):
        return self._im.wcs


class Image:
    """Image [chan, pol, ny, nx] float64 with RASCIL's canonical WCS conventions. The pixels
    may also be a GPU torch tensor (float64 or float32): invert_ng / predict_ng then read
    and write them on the device."""

    def __init__(self, pixels, cellsize, phasecentre, frequency, polarisation_frame,
                 channel_bandwidth=None):
        self.pixels = pixels if (hasattr(pixels, "is_cuda") and pixels.is_cuda) else numpy.asarray(pixels)
        nchan, npol, ny, nx = self.pixels.shape
        assert polarisation_frame.npol == npol
        self.cellsize = float(cellsize)
        self.phasecentre = phasecentre
        self.frequency = numpy.asarray(frequency, "f8").reshape(nchan)
        self.channel_bandwidth = (
            numpy.ones(nchan) * 1e6 if channel_bandwidth is None else numpy.asarray(channel_bandwidth)
        )
        self.polarisation_frame = polarisation_frame
        cdelt_f = self.channel_bandwidth[0] if nchan == 1 else self.frequency[1] - self.frequency[0]
        # image/operations.py:740-751, imaging/base.py:402-418
        self.wcs = _WCS(
            cdelt=[-math.degrees(cellsize), math.degrees(cellsize), 1.0, cdelt_f],
            crpix=[nx // 2 + 1, ny // 2 + 1, 1.0, 1.0],
            crval=[math.degrees(phasecentre.ra), math.degrees(phasecentre.dec), 1.0,
                   self.frequency[0]],
            ctype=["RA---SIN", "DEC--SIN", "STOKES", "FREQ"],
        )

    def __getitem__(self, name):
        return _Field(self, name)

    @property
    def image_acc(self):
        return _ImageAccessor(self)

    def copy(self, deep=True):
        return copy.deepcopy(self) if deep else copy.copy(self)


def image_is_canonical(im: Image) -> bool:
    """4-D, RA---SIN / DEC--SIN / STOKES / FREQ or MOMENT (image/operations.py:75-102)."""
    if im is None:
        return True
    if len(im.pixels.shape) != 4:
        return False
    ct = im.wcs.ctype
    return ct[:3] == ["RA---SIN", "DEC--SIN", "STOKES"] and ct[3] in ("FREQ", "MOMENT")


def create_image(npixel, cellsize, phasecentre, frequency, polarisation_frame=None,
                 channel_bandwidth=None, nchan=None, dtype="float64"):
    """Zero image in RASCIL's canonical layout (image/operations.py:690-759)."""
    if polarisation_frame is None:
        polarisation_frame = PolarisationFrame("stokesI")
    frequency = numpy.atleast_1d(numpy.asarray(frequency, "f8"))
    if nchan is not None:
        frequency = frequency[:nchan]
    pixels = numpy.zeros([len(frequency), polarisation_frame.npol, npixel, npixel], dtype=dtype)
    return Image(pixels, cellsize, phasecentre, frequency, polarisation_frame, channel_bandwidth)


def get_parameter(kwargs, key, default=None):
    """kwargs lookup with default (rascil/data_models/parameters.py:108-125): a key that is
    present is returned as given, None included."""
    if kwargs is None:
        return default
    return kwargs[key] if key in kwargs else default


@dataclass
class Skycomponent:
    """Point source: direction, frequency [nchan], flux [nchan, npol], polarisation frame."""

    direction: Direction
    frequency: numpy.ndarray
    flux: numpy.ndarray
    polarisation_frame: PolarisationFrame
    name: str = ""
    shape: str = "Point"
