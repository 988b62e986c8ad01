"""Gain tables on the GPU: GainTable, create_gaintable_from_blockvisibility, apply_gaintable.

Mirrors rascil/processing_components/calibration/operations.py: apply_gaintable (:37-274),
create_gaintable_from_blockvisibility (:287-360); GainTable follows
rascil/data_models/memory_data_models.py (gain [time, antenna, frequency, nrec, nrec]) and the
receptor frames of rascil/data_models/polarisation.py:32-61. The per-visibility products run
in hipng_apply_gaintable; the host matches visibility times to gain-table rows.
"""

import ctypes

import numpy

from . import _lib
from .data_models import get_parameter

RECEPTORS = {"circular": 2, "circularnp": 2, "linear": 2, "linearnp": 2, "stokesI": 1}


class GainTable:
    """Complex gains [time, antenna, frequency, nrec, nrec] with their weights, the time of each
    row and its interval of validity."""

    def __init__(self, gain, time, interval, weight=None, residual=None, frequency=None,
                 receptor_frame="linear", jones_type="T"):
        self.gain = numpy.asarray(gain, dtype="c16")
        ntimes, nants, nchan, nrec, _ = self.gain.shape
        self.time = numpy.asarray(time, dtype="f8").reshape(ntimes)
        self.interval = numpy.asarray(interval, dtype="f8").reshape(ntimes)
        self.weight = numpy.ones(self.gain.shape) if weight is None else numpy.asarray(weight, "f8")
        self.residual = numpy.zeros((ntimes, nchan, nrec, nrec)) if residual is None else residual
        self.frequency = frequency
        if receptor_frame not in RECEPTORS:
            raise ValueError("Unknown receptor frame %s" % str(receptor_frame))
        self.receptor_frame = receptor_frame
        self.jones_type = jones_type

    def __getitem__(self, name):
        return _Field(self, name)


class _Field:
    def __init__(self, obj, name):
        self._obj, self._name = obj, name

    @property
    def data(self):
        return getattr(self._obj, self._name)

    @data.setter
    def data(self, value):
        setattr(self._obj, self._name, value)


def _nants(vis):
    a1, a2 = vis.baselines
    return int(max(numpy.max(a1), numpy.max(a2))) + 1


def create_gaintable_from_blockvisibility(vis, timeslice=None, frequencyslice=None, jones_type="T",
                                          **kwargs):
    """Unit gains consistent with `vis` (operations.py:287-360): one row per visibility time
    ("auto" / None / <= 0) or per non-empty bin of max(1, ceil(span / timeslice)) equal bins
    (the average time of the bin); one channel for T and G, every channel for B."""
    import pandas

    if timeslice == "auto" or timeslice is None or timeslice <= 0.0:
        utimes = numpy.asarray(vis.time, dtype="f8")
    else:
        t = numpy.asarray(vis.time, dtype="f8")
        nbins = max(1, int(numpy.ceil((numpy.max(t) - numpy.min(t)) / timeslice)))
        cats = pandas.cut(t, nbins)
        utimes = numpy.array([t[cats == c].mean() for c in cats.categories if numpy.any(cats == c)])
    interval = numpy.ones_like(utimes)
    if len(utimes) > 1:
        interval[0] = utimes[1] - utimes[0]
        interval[1:] = utimes[1:] - utimes[:-1]
    if jones_type == "B":
        freq = numpy.asarray(vis.frequency)
    elif jones_type in ("G", "T"):
        freq = numpy.average(vis.frequency) * numpy.ones([1])
    else:
        raise ValueError(f"Unknown Jones type {jones_type}")
    rf = vis.polarisation_frame.type
    if rf not in RECEPTORS:
        raise ValueError("Unknown receptor frame %s" % str(rf))
    nrec = RECEPTORS[rf]
    shape = [len(utimes), _nants(vis), len(freq), nrec, nrec]
    gain = numpy.ones(shape, dtype="complex")
    if nrec > 1:
        gain[..., 0, 1] = 0.0
        gain[..., 1, 0] = 0.0
    return GainTable(gain, utimes, interval, weight=numpy.ones(shape), frequency=freq, receptor_frame=rf,
                     jones_type=jones_type)


def apply_gaintable(vis, gt: GainTable, inverse=False, **kwargs):
    """Apply the gain table to `vis` in place and return it (operations.py:37-274):
    V <- G1 V conj(G2) per baseline (scalar for one polarisation, the diagonal for two), or
    with the inverse gains; keyword use_flags (default False) as the reference."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("hipng gain application needs a ROCm GPU")
    lib = _lib.load()
    ngt, nants, gnchan, nrec, _ = gt.gain.shape
    nt, nbl, nchan, npol = vis.vis.shape
    a1 = numpy.ascontiguousarray(vis.baselines[0], numpy.int32)
    a2 = numpy.ascontiguousarray(vis.baselines[1], numpy.int32)
    if len(a1) != nbl or len(a2) != nbl:
        raise ValueError("baselines do not match the visibility shape")
    if a1.min(initial=0) < 0 or a2.min(initial=0) < 0 or max(a1.max(initial=0), a2.max(initial=0)) >= nants:
        raise ValueError("baseline antenna outside the gain table")
    # visibility times within half a row's interval, rows in order (operations.py:68-73)
    match = numpy.abs(numpy.asarray(vis.time)[None, :] - gt.time[:, None]) < gt.interval[:, None] / 2.0
    npass = int(match.sum(axis=0).max(initial=0))
    if npass == 0:
        return vis
    use_flags = get_parameter(kwargs, "use_flags", False)
    flags = numpy.asarray(vis.flags)
    row_flagged = numpy.array([bool(use_flags and match[r].any() and numpy.max(flags[match[r]]) > 0.0)
                               for r in range(ngt)])
    dev = torch.device("cuda", torch.cuda.current_device())
    st = torch.cuda.current_stream(dev)
    v = torch.from_numpy(numpy.ascontiguousarray(vis.vis, "c16")).to(dev)
    w = torch.from_numpy(numpy.ascontiguousarray(vis.weight, "f8")).to(dev)
    f = torch.from_numpy(numpy.ascontiguousarray(flags, "f8")).to(dev) if row_flagged.any() else None
    g = torch.from_numpy(numpy.ascontiguousarray(gt.gain, "c16")).to(dev)
    gw = torch.from_numpy(numpy.ascontiguousarray(gt.weight, "f8").reshape(gt.gain.shape)).to(dev)
    d1 = torch.from_numpy(a1).to(dev)
    d2 = torch.from_numpy(a2).to(dev)
    order = numpy.cumsum(match, axis=0) * match  # k-th match of a time = row with order k
    for k in range(1, npass + 1):
        t2g = numpy.full(nt, -1, numpy.int32)
        rr, tt = numpy.nonzero(order == k)
        t2g[tt] = rr
        dt2g = torch.from_numpy(t2g).to(dev)
        tf = None
        if f is not None:
            tf = torch.from_numpy(numpy.where(t2g >= 0, row_flagged[numpy.maximum(t2g, 0)], False)
                                  .astype(numpy.int8)).to(dev)
        _lib.check(lib.hipng_apply_gaintable(
            ctypes.c_void_p(v.data_ptr()), ctypes.c_void_p(w.data_ptr()),
            ctypes.c_void_p(f.data_ptr() if f is not None else None),
            ctypes.c_void_p(tf.data_ptr() if tf is not None else None), nt, nbl, nchan, npol,
            ctypes.c_void_p(d1.data_ptr()), ctypes.c_void_p(d2.data_ptr()), ctypes.c_void_p(g.data_ptr()),
            ctypes.c_void_p(gw.data_ptr()), ngt, nants, gnchan, nrec, ctypes.c_void_p(dt2g.data_ptr()),
            1 if inverse else 0, ctypes.c_void_p(st.cuda_stream)))
    vis.vis[...] = v.cpu().numpy()
    vis.weight[...] = w.cpu().numpy()
    return vis
