"""Seeded synthetic observations for the imaging path.

uvw synthesis follows RASCIL: local (east, north, up) station positions rotated by
xyz_at_latitude (rascil/processing_components/util/coordinate_support.py:229-249),
rotated per hour angle by xyz_to_uvw (:352-379), baselines a1 <= a2 with
uvw = pos[a2] - pos[a1] (visibility/base.py:68-79, :258), autocorrelations weighted 0
(:249-256). Observation site: SKA-LOW latitude -26.824722084 deg
(simulation/configurations.py:401-402).
"""

import math

import numpy

from .data_models import BlockVisibility, Direction, PolarisationFrame, Skycomponent

SKA_LOW_LATITUDE = math.radians(-26.824722084)


def xyz_at_latitude(local_xyz, lat):
    x, y, z = numpy.hsplit(local_xyz, 3)
    lat2 = numpy.pi / 2 - lat
    y2 = -z * numpy.sin(lat2) + y * numpy.cos(lat2)
    z2 = z * numpy.cos(lat2) + y * numpy.sin(lat2)
    return numpy.hstack([x, y2, z2])


def xyz_to_uvw(xyz, ha, dec):
    x, y, z = numpy.hsplit(xyz, 3)
    u = x * numpy.cos(ha) - y * numpy.sin(ha)
    v0 = x * numpy.sin(ha) + y * numpy.cos(ha)
    w = z * numpy.sin(dec) - v0 * numpy.cos(dec)
    v = z * numpy.cos(dec) + v0 * numpy.sin(dec)
    return numpy.hstack([u, v, w])


def uvw_to_xyz(uvw, ha, dec):
    """Inverse of xyz_to_uvw (coordinate_support.py:382-407): the transpose of its rotation."""
    u, v, w = numpy.hsplit(uvw, 3)
    v0 = v * numpy.sin(dec) - w * numpy.cos(dec)
    z = v * numpy.cos(dec) + w * numpy.sin(dec)
    x = u * numpy.cos(ha) + v0 * numpy.sin(ha)
    y = -u * numpy.sin(ha) + v0 * numpy.cos(ha)
    return numpy.hstack([x, y, z])


def baseline_indices(nant, autocorrelations=True):
    a1, a2 = numpy.triu_indices(nant, 0 if autocorrelations else 1)
    return a1, a2


def baseline_uvw(ants_enu, ha, dec, latitude=SKA_LOW_LATITUDE, autocorrelations=True):
    """uvw [ntimes, nbaselines, 3] in metres."""
    xyz = xyz_at_latitude(numpy.asarray(ants_enu, "f8"), latitude)
    a1, a2 = baseline_indices(len(xyz), autocorrelations)
    out = numpy.empty((len(ha), len(a1), 3))
    for it, h in enumerate(ha):
        p = xyz_to_uvw(xyz, h, dec)
        out[it] = p[a2] - p[a1]
    return out


def create_blockvisibility(ants_enu, ha, frequency, phasecentre: Direction,
                           polarisation_frame=None, latitude=SKA_LOW_LATITUDE, weight=1.0,
                           autocorrelations=True, channel_bandwidth=None):
    """Zero BlockVisibility for stations `ants_enu` [nant, 3] at hour angles `ha` (radians)."""
    if polarisation_frame is None:
        polarisation_frame = PolarisationFrame("stokesI")
    frequency = numpy.atleast_1d(numpy.asarray(frequency, "f8"))
    uvw = baseline_uvw(ants_enu, ha, phasecentre.dec, latitude, autocorrelations)
    a1, a2 = baseline_indices(len(ants_enu), autocorrelations)
    shape = (len(ha), len(a1), len(frequency), polarisation_frame.npol)
    wt = numpy.full(shape, float(weight))
    wt[:, a1 == a2, ...] = 0.0
    return BlockVisibility(numpy.zeros(shape, "c16"), uvw, frequency, polarisation_frame,
                           phasecentre, weight=wt, imaging_weight=wt.copy(),
                           channel_bandwidth=channel_bandwidth, time=numpy.asarray(ha),
                           baselines=(a1, a2))


def gaussian_disc_stations(n, sigma, rmax, rng):
    """n stations, radius |N(0, sigma)| truncated to rmax, uniform azimuth, zero height."""
    r = numpy.empty(0)
    while r.size < n:
        s = numpy.abs(rng.normal(0.0, sigma, 2 * n))
        r = numpy.concatenate([r, s[s <= rmax]])
    r = r[:n]
    phi = rng.uniform(0.0, 2 * numpy.pi, n)
    return numpy.stack([r * numpy.cos(phi), r * numpy.sin(phi), numpy.zeros(n)], axis=1)


def core_plus_arms_stations(n, core_frac, sigma, rmin, rmax, rng):
    """core_frac Gaussian core (E/N ~ N(0, sigma)), rest log-uniform radius in [rmin, rmax]."""
    nc = int(round(core_frac * n))
    core = rng.normal(0.0, sigma, (nc, 2))
    r = numpy.exp(rng.uniform(numpy.log(rmin), numpy.log(rmax), n - nc))
    phi = rng.uniform(0.0, 2 * numpy.pi, n - nc)
    outer = numpy.stack([r * numpy.cos(phi), r * numpy.sin(phi)], axis=1)
    en = numpy.concatenate([core, outer])
    return numpy.concatenate([en, numpy.zeros((n, 1))], axis=1)


def powerlaw_fluxes(n, index, smin, smax, rng):
    """Inverse-CDF samples of S^index on [smin, smax]."""
    a = index + 1.0
    u = rng.uniform(0.0, 1.0, n)
    return (smin ** a + u * (smax ** a - smin ** a)) ** (1.0 / a)


def config0(seed=0):
    """Plain CPU reference shape: 100 stations, 30 HA in [-3h, 3h], dec -45, 1 ch at 100 MHz."""
    rng = numpy.random.default_rng(seed)
    ants = gaussian_disc_stations(100, 300.0, 1000.0, rng)
    ha = numpy.linspace(-3, 3, 30) * numpy.pi / 12
    pc = Direction(0.0, math.radians(-45.0))
    return create_blockvisibility(ants, ha, [1e8], pc, autocorrelations=False)


def config1_geometry(seed=0, nstations=512, ntimes=128, nchan=64):
    """Production shape geometry: uvw [nrow, 3] and freq [nchan] (no vis allocated)."""
    rng = numpy.random.default_rng(seed)
    ants = core_plus_arms_stations(nstations, 0.7, 500.0, 1000.0, 40000.0, rng)
    ha = numpy.linspace(-2, 2, ntimes) * numpy.pi / 12
    dec = math.radians(-30.0)
    uvw = baseline_uvw(ants, ha, dec, autocorrelations=False).reshape(-1, 3)
    freq = numpy.linspace(100e6, 164e6, nchan, endpoint=False)
    return uvw, freq, Direction(0.0, dec)


def random_point_sources(n, radius, phasecentre, frequency, rng, flux_lo=0.1, flux_hi=1.0,
                         powerlaw=None, pol_frame=None):
    """n point sources uniform within `radius` (rad) of the phase centre."""
    from .data_models import lmn_to_direction

    if pol_frame is None:
        pol_frame = PolarisationFrame("stokesI")
    r = radius * numpy.sqrt(rng.uniform(0, 1, n))
    phi = rng.uniform(0, 2 * numpy.pi, n)
    if powerlaw is None:
        s = rng.uniform(flux_lo, flux_hi, n)
    else:
        s = powerlaw_fluxes(n, powerlaw, flux_lo, flux_hi, rng)
    comps = []
    for k in range(n):
        d = lmn_to_direction(r[k] * numpy.cos(phi[k]), r[k] * numpy.sin(phi[k]), phasecentre)
        flux = numpy.zeros((len(frequency), pol_frame.npol))
        flux[:, 0] = s[k]
        comps.append(Skycomponent(d, numpy.asarray(frequency), flux, pol_frame))
    return comps
