"""ORACLE — test infrastructure only (never imported by rascil_amd).

CPU port (FP64, OpenMP C + multithreaded scipy.fft) of the same w-stacking algorithm the
GPU library runs: bucket sort, per-plane scatter gridding / gather degridding, per-plane
FFT, w-screen and grid correction. It is the step-by-step checker for the GPU kernels and
the CPU baseline of bench.py (ducc0, the reference's own backend, is not installed:
requirements.txt:83 pins ducc0==0.24.0, absent from this image). The plan parameters
(W, beta, grid, tiles, planes) are passed in, so both sides run the same configuration.
"""

import ctypes
import os
import time

import numpy
import scipy.fft
from numpy.polynomial.legendre import leggauss

from . import sortref
from .sortref import OraclePlan, _p

C_M_S = 299792458.0


def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _oplan(plan):
    return OraclePlan(plan["gx"], plan["gy"], plan["W"], plan["tile"], plan["ntx"], plan["nty"],
                      plan["nplanes"], plan["wstacking"], 0, plan["w0"], plan["dw"])


def channel_tables(plan, freq):
    fc = numpy.asarray(freq, "f8") / C_M_S
    sx = fc * plan["pixsize_x"] * float(plan["gx"])
    sy = fc * plan["pixsize_y"] * float(plan["gy"])
    sw = fc / plan["dw"] if plan["wstacking"] else numpy.zeros_like(fc)
    pn = fc * plan["nshift"]
    return sx, sy, sw, pn


def kernel_ft(plan, xi):
    """Phi(xi) = int psi(2s/W) exp(2 pi i s xi) ds by 128-point Gauss-Legendre."""
    W, beta = plan["W"], plan["beta"]
    z, w = leggauss(128)
    keep = z > 0
    z, w = z[keep], w[keep]
    c = W * w * numpy.exp(beta * (numpy.sqrt(1 - z * z) - 1))
    xi = numpy.asarray(xi, "f8")
    return numpy.cos(numpy.pi * W * numpy.multiply.outer(xi, z)) @ c


def correction(plan):
    """1 / (Phi_x Phi_y [Phi_w(dw ((n-1) - nshift)) n]) on the [nx, ny] image."""
    nx, ny = plan["nx"], plan["ny"]
    tx = numpy.arange(nx) - nx // 2
    ty = numpy.arange(ny) - ny // 2
    den = numpy.outer(kernel_ft(plan, tx / plan["gx"]), kernel_ft(plan, ty / plan["gy"]))
    if plan["wstacking"]:
        x = tx * plan["pixsize_x"]
        y = ty * plan["pixsize_y"]
        r2 = x[:, None] ** 2 + y[None, :] ** 2
        rt = numpy.sqrt(1 - r2)
        nm1s = -r2 / (rt + 1) - plan["nshift"]
        den = den * kernel_ft(plan, (plan["dw"] * nm1s).ravel()).reshape(nx, ny) * rt
    return 1.0 / den


class Prepared:
    def __init__(self, uvw, freq, plan, ms=None, wgt=None, nb=1):
        L = sortref.lib()
        self.plan = plan
        self.op = _oplan(plan)
        uvw = numpy.ascontiguousarray(uvw, "f8")
        sx, sy, sw, pn = channel_tables(plan, freq)
        ks, perm = sortref.bucket_sort(uvw, sx, sy, sw, plan)
        n = ks.size
        self.n, self.nb, self.perm, self.nchan = n, nb, perm, len(sx)
        self.ntiles = plan["ntx"] * plan["nty"]
        nkeys = plan["nplanes"] * self.ntiles
        self.bstart = numpy.zeros(nkeys, numpy.uint32)
        self.bend = numpy.zeros(nkeys, numpy.uint32)
        if n:
            first = numpy.flatnonzero(numpy.r_[True, ks[1:] != ks[:-1]])
            self.bstart[ks[first]] = first
            self.bend[ks[first]] = numpy.r_[first[1:], n]
        self.splane = (ks // self.ntiles).astype(numpy.int32)
        self.ukeys = ks[first] if n else numpy.zeros(0, numpy.uint32)
        self.sstart = numpy.searchsorted(ks, numpy.arange(plan["nplanes"] + 1) * self.ntiles)
        self.ax = numpy.empty(n, numpy.int32)
        self.ay = numpy.empty(n, numpy.int32)
        self.f3 = numpy.empty((n, 3))
        self.vals = None
        msv = wv = None
        if ms is not None:
            msv = numpy.ascontiguousarray(ms, "c16").reshape(-1)
            self.vals = numpy.empty(n * nb, "c16")
            if wgt is not None:
                wv = numpy.ascontiguousarray(wgt, "f8").reshape(-1)
        L.oracle_records(_p(uvw), ctypes.c_int32(self.nchan), _p(perm), ctypes.c_int64(n), _p(sx),
                         _p(sy), _p(sw), _p(pn), ctypes.byref(self.op),
                         _p(msv) if msv is not None else None, _p(wv) if wv is not None else None,
                         ctypes.c_int32(nb), _p(self.ax), _p(self.ay), _p(self.f3),
                         _p(self.vals) if self.vals is not None else None)
        self.uvw, self.pn = uvw, pn

    def grid_plane(self, p, plane):
        P = self.plan
        ww = P["W"] if P["wstacking"] else 1
        lo = max(0, p - ww + 1) * self.ntiles
        hi = (min(p, P["nplanes"] - ww) + 1) * self.ntiles
        sel = self.ukeys[(self.ukeys >= lo) & (self.ukeys < hi)]
        tiles = numpy.ascontiguousarray(numpy.unique(sel % self.ntiles), dtype=numpy.uint32)
        sortref.lib().oracle_grid_plane(ctypes.c_int32(p), ctypes.byref(self.op),
                                        ctypes.c_double(P["beta"]), _p(self.bstart),
                                        _p(self.bend), _p(self.ax), _p(self.ay), _p(self.f3),
                                        _p(self.vals), ctypes.c_int32(self.nb), _p(tiles),
                                        ctypes.c_int64(tiles.size), _p(plane))

    def degrid_plane(self, p, plane, acc):
        P = self.plan
        ww = P["W"] if P["wstacking"] else 1
        j0 = self.sstart[max(0, p - ww + 1)]
        j1 = self.sstart[min(p, P["nplanes"] - ww) + 1]
        sortref.lib().oracle_degrid_plane(ctypes.c_int32(p), ctypes.byref(self.op),
                                          ctypes.c_double(P["beta"]), ctypes.c_int64(int(j0)),
                                          ctypes.c_int64(int(j1)), _p(self.splane), _p(self.ax),
                                          _p(self.ay), _p(self.f3), _p(plane), ctypes.c_int32(self.nb),
                                          _p(acc))


def _screen_acc(prep, wp, plane, acc):
    P = prep.plan
    sortref.lib().oracle_screen_acc(ctypes.byref(prep.op), ctypes.c_int32(P["nx"]), ctypes.c_int32(P["ny"]),
                                    ctypes.c_double(P["pixsize_x"]), ctypes.c_double(P["pixsize_y"]),
                                    ctypes.c_double(P["nshift"]), ctypes.c_double(wp), _p(plane),
                                    ctypes.c_int32(prep.nb), _p(acc))


def _screen_fill(prep, wp, dc, plane):
    P = prep.plan
    sortref.lib().oracle_screen_fill(ctypes.byref(prep.op), ctypes.c_int32(P["nx"]), ctypes.c_int32(P["ny"]),
                                     ctypes.c_double(P["pixsize_x"]), ctypes.c_double(P["pixsize_y"]),
                                     ctypes.c_double(P["nshift"]), ctypes.c_double(wp), _p(dc),
                                     ctypes.c_int32(prep.nb), _p(plane))


def cpu_ms2dirty(uvw, freq, ms, wgt, plan):
    """ms [nrow, nchan, nb] -> dirty [nb, nx, ny] with the GPU's plan parameters (FP64)."""
    ms = numpy.asarray(ms)
    nb = ms.shape[2] if ms.ndim == 3 else 1
    prep = Prepared(uvw, freq, plan, ms, wgt, nb)
    P = plan
    acc = numpy.zeros((nb, P["nx"], P["ny"]))
    plane = numpy.zeros((nb, P["gx"], P["gy"]), "c16")
    for p in range(P["nplanes"]):
        plane[...] = 0
        prep.grid_plane(p, plane)
        F = scipy.fft.ifft2(plane, axes=(1, 2), norm="forward", workers=_threads())
        _screen_acc(prep, P["w0"] + p * P["dw"], numpy.ascontiguousarray(F), acc)
    return acc * correction(P)[None]


def cpu_dirty2ms(uvw, freq, dirty, wgt, plan):
    """dirty [nb, nx, ny] -> ms [nrow, nchan, nb] with the GPU's plan parameters (FP64)."""
    dirty = numpy.asarray(dirty, "f8")
    if dirty.ndim == 2:
        dirty = dirty[None]
    nb = dirty.shape[0]
    prep = Prepared(uvw, freq, plan, None, None, nb)
    P = plan
    dc = numpy.ascontiguousarray(dirty * correction(P)[None])
    acc = numpy.zeros(prep.n * nb, "c16")
    plane = numpy.zeros((nb, P["gx"], P["gy"]), "c16")
    for p in range(P["nplanes"]):
        _screen_fill(prep, P["w0"] + p * P["dw"], dc, plane)
        F = numpy.ascontiguousarray(scipy.fft.fft2(plane, axes=(1, 2), workers=_threads()))
        prep.degrid_plane(p, F, acc)
    return _unsort(prep, acc, uvw, wgt)


def _unsort(prep, acc, uvw, wgt):
    nb = prep.nb
    perm = prep.perm.astype(numpy.int64)
    rows = perm // prep.nchan
    chans = perm % prep.nchan
    w = uvw[rows, 2]
    t = numpy.abs(w) * prep.pn[chans]
    ph = numpy.exp(2j * numpy.pi * (t - numpy.rint(t))) if prep.plan["wstacking"] else 1.0
    a = acc.reshape(prep.n, nb) * (ph[:, None] if prep.plan["wstacking"] else 1.0)
    a = numpy.where((w < 0)[:, None], a.conj(), a)
    out = numpy.empty_like(a)
    out[perm] = a
    out = out.reshape(uvw.shape[0], prep.nchan, nb)
    if wgt is not None:
        out = out * numpy.asarray(wgt, "f8").reshape(out.shape)
    return out


def _time_rows(uvw, freq, npol, P, Pp, rows):
    """Seconds to sort, grid and degrid the given rows through every plane (CPU port)."""
    rng = numpy.random.default_rng(0)
    su = uvw[rows]
    nchan = len(freq)
    ms = rng.normal(size=(len(rows), nchan, npol)) + 1j * rng.normal(size=(len(rows), nchan, npol))
    t0 = time.time()
    prep = Prepared(su, freq, P, ms, numpy.ones((len(rows), nchan, npol)), npol)
    plane = numpy.zeros((npol, P["gx"], P["gy"]), "c16")
    for p in range(P["nplanes"]):
        prep.grid_plane(p, plane)
    prep2 = Prepared(su, freq, Pp, None, None, npol)
    acc = numpy.zeros(prep2.n * npol, "c16")
    for p in range(Pp["nplanes"]):
        prep2.degrid_plane(p, plane, acc)
    _unsort(prep2, acc, su, None)
    return time.time() - t0, plane


def _time_planes(P, npol, plane, nthreads):
    """Seconds for one full plane of each direction (zero/fill, FFT, screen)."""
    prep = Prepared(numpy.zeros((1, 3)), [1e8], P, numpy.zeros((1, 1, npol), "c16"), None, npol)
    acc_img = numpy.zeros((npol, P["nx"], P["ny"]))
    t0 = time.time()
    plane[...] = 0
    F = scipy.fft.ifft2(plane, axes=(1, 2), norm="forward", workers=nthreads)
    _screen_acc(prep, P["w0"], numpy.ascontiguousarray(F), acc_img)
    t_inv = time.time() - t0
    del F
    dc = numpy.zeros((npol, P["nx"], P["ny"]))
    t0 = time.time()
    _screen_fill(prep, P["w0"], dc, plane)
    F = scipy.fft.fft2(plane, axes=(1, 2), workers=nthreads)
    t_fwd = time.time() - t0
    return t_inv, t_fwd


def cpu_baseline(wl, plan_i, plan_p, seconds=20.0, plans=None, frac=0.01):
    """Time the CPU port on a bounded sample of the bench workload and extrapolate.

    The candidate plans (the GPU's and, when given, other oversampling factors; a CPU
    gridder would pick its own parameters) are compared on a small row sample; the fastest
    one is then timed on `frac` / 2 and `frac` of the rows (sort, grid and degrid through
    every plane), and the row-proportional time is the line through those two points
    extended to all rows. The per-plane part (one full plane of each direction: fill, FFT,
    screen) is timed once and multiplied by the plane count. `seconds` bounds the small
    sample (rows x channels x polarisations of about 1.5e5 visibilities each)."""
    nthreads = _threads()
    uvw, freq, npol = wl["uvw"], wl["freq"], wl["npol"]
    nrow, nchan = uvw.shape[0], len(freq)
    rng = numpy.random.default_rng(0)
    perm = rng.permutation(nrow)
    rs = max(2, min(nrow, int(1.5e5 / (nchan * npol)) + 1))
    cands = [(plan_i, plan_p)] + list(plans or [])
    best = None
    for P, Pp in cands:
        ts, plane = _time_rows(uvw, freq, npol, P, Pp, numpy.sort(perm[:rs]))
        t_inv, t_fwd = _time_planes(P, npol, plane, nthreads)
        del plane
        est = ts * nrow / rs + t_inv * P["nplanes"] + t_fwd * Pp["nplanes"]
        if best is None or est < best[0]:
            best = (est, P, Pp, t_inv, t_fwd)
    _, P, Pp, t_inv, t_fwd = best
    r2 = max(2, min(nrow, int(round(frac * nrow))))
    r1 = max(1, r2 // 2)
    t1, _ = _time_rows(uvw, freq, npol, P, Pp, numpy.sort(perm[:r1]))
    t2, plane = _time_rows(uvw, freq, npol, P, Pp, numpy.sort(perm[:r2]))
    del plane
    slope = max(0.0, (t2 - t1) / (r2 - r1))
    t_rows = t2 + slope * (nrow - r2)
    total = t_rows + t_inv * P["nplanes"] + t_fwd * Pp["nplanes"]
    nvis = nrow * nchan * npol
    desc = (f"CPU port of the same algorithm (FP64, OpenMP C + scipy.fft; ducc0 is not installed), "
            f"fastest of {len(cands)} plans on a {rs}-row sample: sigma {P['sigma']} W {P['W']} grid {P['gx']} "
            f"planes {P['nplanes']}; rows {r1} and {r2} ({100.0 * r2 / nrow:.2f} % of {nrow}) x {nchan} ch x "
            f"{npol} pol through all planes ({t1:.1f} / {t2:.1f} s, linear fit -> {t_rows:.0f} s) + one "
            f"{P['gx']}^2 x {npol} plane per direction ({t_inv:.1f} + {t_fwd:.1f} s, x{P['nplanes']}) = {total:.0f} s")
    return {
        "value": round(nvis / total / 1e6, 4), "unit": "Mvis/s", "cores": nthreads, "kind": "port",
        "sample": desc, "extrapolated_seconds": round(total, 1),
        "fit": {"rows": [int(r1), int(r2)], "seconds": [round(t1, 2), round(t2, 2)],
                "fraction_of_rows": round(r2 / nrow, 4), "plane_seconds": [round(t_inv, 2), round(t_fwd, 2)]},
    }
