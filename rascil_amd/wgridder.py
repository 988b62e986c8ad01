"""ducc0.wgridder-shaped backend (ms2dirty / dirty2ms) running on the hipng HIP library.

Same calling surface as ducc0 0.24 ``wgridder.ms2dirty`` / ``wgridder.dirty2ms`` as RASCIL
uses them (rascil/processing_components/imaging/ng.py:91-103, 232-248): positional
uvw, freq, ms/dirty, wgt, npix_x, npix_y, pixsize_x, pixsize_y, nu, nv, epsilon,
do_wstacking, nthreads; keyword verbosity, mask, double_precision_accumulation.

Inputs may be numpy arrays (copied to the current GPU, result returned as numpy, like
ducc0) or torch tensors already on the GPU (result returned as a torch tensor on the
same device, computed on the current stream).

Extensions: ms of shape [nrow, nchan, nbatch] (and dirty [nbatch, npix_x, npix_y])
processes up to nbatch independent images sharing uvw/freq in one pass; `precision`
selects "fp64" (default for complex128 input), "mixed" (default for complex64 input) or
"fp32"; `plan_out` (a dict) receives the kernel/grid/plane parameters that were used;
`plan_opts` (a dict, see _lib.plan_opts) pins plan choices (oversampling, support, tile,
dense-run threshold, task sizes, uv extent) without touching the process environment.
"""

import ctypes

import numpy

from . import _lib

MAX_BATCH = 4


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("hipng needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch


def _precision(precision, dtype_is_double):
    if precision is None:
        precision = "fp64" if dtype_is_double else "mixed"
    if precision not in _lib.PRECISIONS:
        raise ValueError(f"unknown precision {precision!r}; use one of {list(_lib.PRECISIONS)}")
    return precision


def _cdtype(torch, precision):
    return torch.complex128 if precision == "fp64" else torch.complex64


def _rdtype(torch, precision):
    return torch.float64 if precision == "fp64" else torch.float32


def _freq_host(freq):
    if hasattr(freq, "detach"):
        freq = freq.detach().cpu().numpy()
    return numpy.ascontiguousarray(numpy.atleast_1d(freq), dtype=numpy.float64)


def _dev(torch, x, dtype, device):
    if isinstance(x, numpy.ndarray) or not hasattr(x, "detach"):
        x = torch.from_numpy(numpy.ascontiguousarray(x))
    return x.to(device=device, dtype=dtype).contiguous()


def _stream(torch, device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _split_batch(fn, kwargs):
    """Retry with half the batch when the plan reports that the batch does not fit in LDS."""
    bc = kwargs.get("batch_chunk") or MAX_BATCH
    while True:
        try:
            return fn(**dict(kwargs, batch_chunk=bc))
        except ValueError as e:
            if "split the batch" not in str(e) or bc == 1:
                raise
            bc //= 2


def ms2dirty(uvw, freq, ms, wgt, npix_x, npix_y, pixsize_x, pixsize_y, nu=0, nv=0,
             epsilon=1e-7, do_wstacking=False, nthreads=1, verbosity=0, mask=None,
             double_precision_accumulation=False, precision=None, plan_out=None, batch_chunk=None,
             plan_opts=None):
    """Weighted visibilities -> dirty image(s); see module docstring.

    Returns [npix_x, npix_y] (or [nbatch, npix_x, npix_y] for batched ms); float64 for
    "fp64", float32 otherwise.
    """
    return _split_batch(_ms2dirty, dict(
        uvw=uvw, freq=freq, ms=ms, wgt=wgt, npix_x=npix_x, npix_y=npix_y, pixsize_x=pixsize_x,
        pixsize_y=pixsize_y, epsilon=epsilon, do_wstacking=do_wstacking, verbosity=verbosity,
        mask=mask, precision=precision, plan_out=plan_out, batch_chunk=batch_chunk, plan_opts=plan_opts))


def dirty2ms(uvw, freq, dirty, wgt, pixsize_x, pixsize_y, nu=0, nv=0, epsilon=1e-7,
             do_wstacking=False, nthreads=1, verbosity=0, mask=None, precision=None,
             plan_out=None, batch_chunk=None, plan_opts=None):
    """Dirty image(s) -> visibilities [nrow, nchan] (or [nrow, nchan, nbatch])."""
    return _split_batch(_dirty2ms, dict(
        uvw=uvw, freq=freq, dirty=dirty, wgt=wgt, pixsize_x=pixsize_x, pixsize_y=pixsize_y,
        epsilon=epsilon, do_wstacking=do_wstacking, verbosity=verbosity, mask=mask,
        precision=precision, plan_out=plan_out, batch_chunk=batch_chunk, plan_opts=plan_opts))


def _ms2dirty(uvw, freq, ms, wgt, npix_x, npix_y, pixsize_x, pixsize_y, nu=0, nv=0,
              epsilon=1e-7, do_wstacking=False, nthreads=1, verbosity=0, mask=None,
              double_precision_accumulation=False, precision=None, plan_out=None, batch_chunk=None,
              plan_opts=None):
    torch = _torch()
    lib = _lib.load()
    as_numpy = isinstance(ms, numpy.ndarray)
    is_double = (ms.dtype == numpy.complex128) if as_numpy else (ms.dtype == torch.complex128)
    precision = _precision(precision, is_double)
    device = torch.device("cuda", torch.cuda.current_device()) if as_numpy or not ms.is_cuda else ms.device
    freq_h = _freq_host(freq)
    nchan = freq_h.shape[0]
    uvw_d = _dev(torch, uvw, torch.float64, device)
    nrow = uvw_d.shape[0]
    if tuple(uvw_d.shape) != (nrow, 3):
        raise ValueError("uvw must have shape [nrow, 3]")
    ms_d = _dev(torch, ms, _cdtype(torch, precision), device)
    batched = ms_d.dim() == 3
    if not batched:
        ms_d = ms_d.reshape(nrow, nchan, 1)
    if tuple(ms_d.shape[:2]) != (nrow, nchan):
        raise ValueError(f"ms shape {tuple(ms_d.shape)} does not match [nrow={nrow}, nchan={nchan}]")
    nb_total = ms_d.shape[2]
    wgt_d = None
    if wgt is not None:
        wgt_d = _dev(torch, wgt, _rdtype(torch, precision), device).reshape(nrow, nchan, -1)
    if mask is not None:
        m = _dev(torch, mask, _rdtype(torch, precision), device).reshape(nrow, nchan, -1)
        wgt_d = m if wgt_d is None else wgt_d * m
    if wgt_d is not None and wgt_d.shape[2] != nb_total:
        wgt_d = wgt_d.expand(nrow, nchan, nb_total).contiguous()
    out = torch.empty((nb_total, npix_x, npix_y), dtype=_rdtype(torch, precision), device=device)
    plan = _lib.HipngPlan()
    bc = min(MAX_BATCH, batch_chunk or MAX_BATCH)
    for b0 in range(0, nb_total, bc):
        b1 = min(nb_total, b0 + bc)
        msb = ms_d[:, :, b0:b1].contiguous() if nb_total > bc else ms_d
        wb = None
        if wgt_d is not None:
            wb = wgt_d[:, :, b0:b1].contiguous() if nb_total > bc else wgt_d
        with torch.cuda.device(device):
            code = lib.hipng_ms2dirty(
                ctypes.c_void_p(uvw_d.data_ptr()), _lib.dptr(freq_h), nrow, nchan,
                ctypes.c_void_p(msb.data_ptr()),
                ctypes.c_void_p(wb.data_ptr()) if wb is not None else None,
                b1 - b0, int(npix_x), int(npix_y), float(pixsize_x), float(pixsize_y),
                float(epsilon), int(bool(do_wstacking)), _lib.PRECISIONS[precision],
                ctypes.c_void_p(out[b0:b1].data_ptr()), _stream(torch, device), _lib.plan_opts(plan_opts),
                ctypes.byref(plan))
        _lib.check(code)
    if plan_out is not None:
        plan_out.update(plan.as_dict())
    if verbosity:
        print("hipng ms2dirty plan:", plan.as_dict())
    if not batched:
        out = out[0]
    if as_numpy:
        return out.cpu().numpy()
    return out


def _dirty2ms(uvw, freq, dirty, wgt, pixsize_x, pixsize_y, nu=0, nv=0, epsilon=1e-7,
              do_wstacking=False, nthreads=1, verbosity=0, mask=None, precision=None,
              plan_out=None, batch_chunk=None, plan_opts=None):
    torch = _torch()
    lib = _lib.load()
    as_numpy = isinstance(dirty, numpy.ndarray)
    is_double = (dirty.dtype == numpy.float64) if as_numpy else (dirty.dtype == torch.float64)
    precision = _precision(precision, is_double)
    device = (torch.device("cuda", torch.cuda.current_device())
              if as_numpy or not dirty.is_cuda else dirty.device)
    freq_h = _freq_host(freq)
    nchan = freq_h.shape[0]
    uvw_d = _dev(torch, uvw, torch.float64, device)
    nrow = uvw_d.shape[0]
    if tuple(uvw_d.shape) != (nrow, 3):
        raise ValueError("uvw must have shape [nrow, 3]")
    d = _dev(torch, dirty, _rdtype(torch, precision), device)
    batched = d.dim() == 3
    if not batched:
        d = d.reshape(1, *d.shape)
    nb_total, npix_x, npix_y = d.shape
    wgt_d = None
    if wgt is not None:
        wgt_d = _dev(torch, wgt, _rdtype(torch, precision), device).reshape(nrow, nchan, -1)
    if mask is not None:
        m = _dev(torch, mask, _rdtype(torch, precision), device).reshape(nrow, nchan, -1)
        wgt_d = m if wgt_d is None else wgt_d * m
    if wgt_d is not None and wgt_d.shape[2] != nb_total:
        wgt_d = wgt_d.expand(nrow, nchan, nb_total).contiguous()
    out = torch.empty((nrow, nchan, nb_total), dtype=_cdtype(torch, precision), device=device)
    plan = _lib.HipngPlan()
    bc = min(MAX_BATCH, batch_chunk or MAX_BATCH)
    for b0 in range(0, nb_total, bc):
        b1 = min(nb_total, b0 + bc)
        chunked = nb_total > bc
        ob = torch.empty((nrow, nchan, b1 - b0), dtype=out.dtype, device=device) if chunked else out
        wb = None
        if wgt_d is not None:
            wb = wgt_d[:, :, b0:b1].contiguous() if chunked else wgt_d
        db = d[b0:b1].contiguous()
        with torch.cuda.device(device):
            code = lib.hipng_dirty2ms(
                ctypes.c_void_p(uvw_d.data_ptr()), _lib.dptr(freq_h), nrow, nchan,
                ctypes.c_void_p(db.data_ptr()),
                ctypes.c_void_p(wb.data_ptr()) if wb is not None else None,
                b1 - b0, int(npix_x), int(npix_y), float(pixsize_x), float(pixsize_y),
                float(epsilon), int(bool(do_wstacking)), _lib.PRECISIONS[precision],
                ctypes.c_void_p(ob.data_ptr()), _stream(torch, device), _lib.plan_opts(plan_opts),
                ctypes.byref(plan))
        _lib.check(code)
        if chunked:
            out[:, :, b0:b1] = ob
    if plan_out is not None:
        plan_out.update(plan.as_dict())
    if verbosity:
        print("hipng dirty2ms plan:", plan.as_dict())
    if not batched:
        out = out[:, :, 0]
    if as_numpy:
        return out.cpu().numpy()
    return out


def plan_params(umax, vmax, wmin, wmax, npix_x, npix_y, pixsize_x, pixsize_y, epsilon,
                do_wstacking, precision="fp64", nbatch=1, nvis=0, plan_opts=None):
    """Host-only plan choice (no GPU needed); returns the hipng_plan structure."""
    lib = _lib.load()
    plan = _lib.HipngPlan()
    _lib.check(lib.hipng_plan_params(umax, vmax, wmin, wmax, npix_x, npix_y, pixsize_x, pixsize_y,
                                     epsilon, int(bool(do_wstacking)), _lib.PRECISIONS[precision],
                                     nbatch, int(nvis), _lib.plan_opts(plan_opts), ctypes.byref(plan)))
    return plan


def channel_scales(plan, freq):
    """Per-channel coordinate scales of the bucket-sort contract (sx, sy, sw)."""
    lib = _lib.load()
    f = _freq_host(freq)
    sx, sy, sw = (numpy.zeros_like(f) for _ in range(3))
    _lib.check(lib.hipng_channel_scales(ctypes.byref(plan), _lib.dptr(f), f.shape[0],
                                        _lib.dptr(sx), _lib.dptr(sy), _lib.dptr(sw)))
    return sx, sy, sw


def bucket_sort(uvw, freq, plan, plan_opts=None):
    """GPU bucket sort by (w plane, uv tile): returns (keys, perm) as numpy uint32.
    plan_opts {"sort_pack": False} recomputes the bucket keys after the digit passes."""
    torch = _torch()
    lib = _lib.load()
    f = _freq_host(freq)
    device = torch.device("cuda", torch.cuda.current_device())
    uvw_d = _dev(torch, uvw, torch.float64, device)
    n = uvw_d.shape[0] * f.shape[0]
    keys = torch.empty(n, dtype=torch.int32, device=device)
    perm = torch.empty(n, dtype=torch.int32, device=device)
    _lib.check(lib.hipng_bucket_sort(ctypes.c_void_p(uvw_d.data_ptr()), _lib.dptr(f), uvw_d.shape[0],
                                     f.shape[0], ctypes.byref(plan), _lib.plan_opts(plan_opts),
                                     ctypes.c_void_p(keys.data_ptr()),
                                     ctypes.c_void_p(perm.data_ptr()), _stream(torch, device)))
    torch.cuda.synchronize(device)
    return keys.cpu().numpy().view(numpy.uint32), perm.cpu().numpy().view(numpy.uint32)


def uvw_extent(uvw, freq):
    """(max|u|, max|v|, min|w|, max|w|) in wavelengths, computed on the GPU."""
    torch = _torch()
    lib = _lib.load()
    f = _freq_host(freq)
    device = torch.device("cuda", torch.cuda.current_device())
    uvw_d = _dev(torch, uvw, torch.float64, device)
    out = numpy.zeros(4)
    _lib.check(lib.hipng_uvw_extent(ctypes.c_void_p(uvw_d.data_ptr()), uvw_d.shape[0], _lib.dptr(f),
                                    f.shape[0], _lib.dptr(out), _stream(torch, device)))
    return out


class Geometry:
    """Plan once, execute many: a fixed visibility geometry (uvw, freq), image and accuracy.

    Construction runs the extent reduction, plan choice, bucket sort and coordinate records
    once and allocates every device buffer. ms2dirty / dirty2ms then only launch kernels on
    the current stream (no allocation, no host synchronisation), so they can be captured in
    a torch.cuda.CUDAGraph (hipGraph) and replayed. Results equal the one-shot functions.
    """

    def __init__(self, uvw, freq, npix_x, npix_y, pixsize_x, pixsize_y, epsilon,
                 do_wstacking=True, precision="mixed", nbatch=1, device=None, plan_opts=None):
        torch = _torch()
        self._lib = _lib.load()
        if precision not in _lib.PRECISIONS:
            raise ValueError(f"unknown precision {precision!r}")
        if not 1 <= nbatch <= MAX_BATCH:
            raise ValueError(f"nbatch must be in [1, {MAX_BATCH}]")
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.freq = _freq_host(freq)
        self.nchan = self.freq.shape[0]
        uvw_d = _dev(torch, uvw, torch.float64, self.device)
        self.nrow = uvw_d.shape[0]
        self.npix_x, self.npix_y, self.nbatch, self.precision = int(npix_x), int(npix_y), nbatch, precision
        self.cdtype, self.rdtype = _cdtype(torch, precision), _rdtype(torch, precision)
        self._handle = ctypes.c_void_p()
        plan = _lib.HipngPlan()
        with torch.cuda.device(self.device):
            _lib.check(self._lib.hipng_geometry_create(
                ctypes.c_void_p(uvw_d.data_ptr()), _lib.dptr(self.freq), self.nrow, self.nchan, nbatch,
                self.npix_x, self.npix_y, float(pixsize_x), float(pixsize_y), float(epsilon),
                int(bool(do_wstacking)), _lib.PRECISIONS[precision], _stream(torch, self.device),
                _lib.plan_opts(plan_opts), ctypes.byref(self._handle), ctypes.byref(plan)))
        self.plan = plan.as_dict()

    def _shape_ms(self):
        return (self.nrow, self.nchan, self.nbatch)

    def _check(self, name, t, dtype, shape):
        """t must be a contiguous `dtype` tensor of `shape` on this geometry's device (its raw
        pointer goes to the kernels)."""
        if (t.dtype != dtype or not t.is_contiguous() or t.device != self.device
                or tuple(t.shape) != tuple(shape)):
            raise ValueError("%s must be a contiguous %s tensor of shape %s on %s (got %s %s on %s)"
                             % (name, dtype, tuple(shape), self.device, t.dtype, tuple(t.shape), t.device))

    def _wgt(self, wgt):
        if wgt is None:
            return None
        if wgt.numel() != self.nrow * self.nchan * self.nbatch:
            raise ValueError("wgt must have %d elements" % (self.nrow * self.nchan * self.nbatch))
        wgt = wgt.reshape(self._shape_ms())
        self._check("wgt", wgt, self.rdtype, self._shape_ms())
        return wgt

    def ms2dirty(self, ms, wgt=None, out=None):
        """ms [nrow, nchan, nbatch] (or [nrow, nchan] when nbatch == 1) GPU tensor -> dirty."""
        torch = _torch()
        if ms.numel() != self.nrow * self.nchan * self.nbatch:
            raise ValueError("ms must have %d elements" % (self.nrow * self.nchan * self.nbatch))
        ms = ms.reshape(self._shape_ms())
        self._check("ms", ms, self.cdtype, self._shape_ms())
        wgt = self._wgt(wgt)
        if out is None:
            out = torch.empty((self.nbatch, self.npix_x, self.npix_y), dtype=self.rdtype, device=self.device)
        else:
            self._check("out", out, self.rdtype, (self.nbatch, self.npix_x, self.npix_y))
        _lib.check(self._lib.hipng_geometry_ms2dirty(
            self._handle, ctypes.c_void_p(ms.data_ptr()),
            ctypes.c_void_p(wgt.data_ptr()) if wgt is not None else None,
            ctypes.c_void_p(out.data_ptr()), _stream(torch, self.device)))
        return out

    def dirty2ms(self, dirty, wgt=None, out=None):
        """dirty [nbatch, npix_x, npix_y] (or [npix_x, npix_y]) GPU tensor -> ms [nrow, nchan, nbatch]."""
        torch = _torch()
        if dirty.numel() != self.nbatch * self.npix_x * self.npix_y:
            raise ValueError("dirty must have %d elements" % (self.nbatch * self.npix_x * self.npix_y))
        dirty = dirty.reshape(self.nbatch, self.npix_x, self.npix_y)
        self._check("dirty", dirty, self.rdtype, (self.nbatch, self.npix_x, self.npix_y))
        wgt = self._wgt(wgt)
        if out is None:
            out = torch.empty(self._shape_ms(), dtype=self.cdtype, device=self.device)
        else:
            if out.numel() != self.nrow * self.nchan * self.nbatch:
                raise ValueError("out must have %d elements" % (self.nrow * self.nchan * self.nbatch))
            self._check("out", out.reshape(self._shape_ms()) if out.is_contiguous() else out, self.cdtype,
                        self._shape_ms())
        _lib.check(self._lib.hipng_geometry_dirty2ms(
            self._handle, ctypes.c_void_p(dirty.data_ptr()),
            ctypes.c_void_p(wgt.data_ptr()) if wgt is not None else None,
            ctypes.c_void_p(out.data_ptr()), _stream(torch, self.device)))
        return out

    def invert_columns(self, cols, vis, weight=None, flags=None, out=None):
        """Invert straight from BlockVisibility-layout columns (hipng_geometry_invert_columns):
        vis [nrow, nchan_cols, npol] (complex128 / complex64), weight (float64 / float32) and
        flags (int32 / int8) of the same shape or None; `cols` from make_columns. Returns the
        dirty images [nbatch, ...] in the order and type `cols` names."""
        torch = _torch()
        shape = (self.nrow, cols.nchan, cols.npol)
        self._check("vis", vis.reshape(shape) if vis.is_contiguous() else vis,
                    {0: torch.complex128, 1: torch.complex64}[cols.vis_type], shape)
        if weight is not None:
            self._check("weight", weight.reshape(shape) if weight.is_contiguous() else weight,
                        {0: torch.float64, 1: torch.float32}[cols.weight_type], shape)
        if flags is not None:
            self._check("flags", flags.reshape(shape) if flags.is_contiguous() else flags,
                        {0: torch.int32, 1: torch.int8}[cols.flag_type], shape)
        oshape = self._image_shape(cols)
        odt = {0: torch.float64, 1: torch.float32}[cols.image_type]
        if out is None:
            out = torch.empty(oshape, dtype=odt, device=self.device)
        else:
            self._check("out", out, odt, oshape)
        _lib.check(self._lib.hipng_geometry_invert_columns(
            self._handle, ctypes.byref(cols), ctypes.c_void_p(vis.data_ptr()),
            ctypes.c_void_p(weight.data_ptr()) if weight is not None else None,
            ctypes.c_void_p(flags.data_ptr()) if flags is not None else None,
            ctypes.c_void_p(out.data_ptr()), _stream(torch, self.device)))
        return out

    def predict_columns(self, cols, dirty, vis_out, model_out=None):
        """Predict into BlockVisibility-layout columns (hipng_geometry_predict_columns): writes
        the geometry's channels of vis_out [nrow, nchan_cols, npol] (and of model_out, the
        visibilities before the gains); dirty [nbatch, ...] in the order and type `cols` names."""
        torch = _torch()
        shape = (self.nrow, cols.nchan, cols.npol)
        vdt = {0: torch.complex128, 1: torch.complex64}[cols.vis_type]
        self._check("vis_out", vis_out.reshape(shape) if vis_out.is_contiguous() else vis_out, vdt, shape)
        if model_out is not None:
            self._check("model_out", model_out.reshape(shape) if model_out.is_contiguous() else model_out, vdt,
                        shape)
        self._check("dirty", dirty, {0: torch.float64, 1: torch.float32}[cols.image_type], self._image_shape(cols))
        _lib.check(self._lib.hipng_geometry_predict_columns(
            self._handle, ctypes.byref(cols), ctypes.c_void_p(dirty.data_ptr()), ctypes.c_void_p(vis_out.data_ptr()),
            ctypes.c_void_p(model_out.data_ptr()) if model_out is not None else None, _stream(torch, self.device)))
        return vis_out

    def _image_shape(self, cols):
        if cols.image_order == 1:
            return (self.nbatch, self.npix_y, self.npix_x)
        return (self.nbatch, self.npix_x, self.npix_y)

    def stats(self):
        """Diagnostics: visibilities, dense-run tasks, visibilities on the dense-run path,
        whether that path is enabled for this geometry, and whether its out-of-place FFT
        plans were found to leave their inputs untouched."""
        out = (ctypes.c_int64 * 5)()
        _lib.check(self._lib.hipng_geometry_stats(self._handle, out, 5))
        return {"nvis": out[0], "dense_tasks": out[1], "dense_vis": out[2], "dense_enabled": bool(out[3]),
                "fft_keeps_input": bool(out[4])}

    def close(self):
        if getattr(self, "_handle", None) and self._handle.value:
            self._lib.hipng_geometry_destroy(self._handle)
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_columns(nchan, npol, conv, chan0=0, vis_dtype=None, weight_dtype=None, flag_dtype=None, lmn=None,
                 gains=None, ant1=None, ant2=None, psf=False, image_order=0, image_dtype=None, scale=None):
    """hipng_columns for Geometry.invert_columns / predict_columns.

    nchan, npol: the columns' channels and polarisations; conv: complex matrix [npol_out,
    npol_in] (columns -> image for invert, image -> columns for predict); chan0: the columns'
    channel of the geometry's first channel; dtypes: torch dtypes of the column arrays
    (complex128 / complex64, float64 / float32, int32 / int8) and of the images
    (float64 / float32); lmn: coefficients (L, M, N) of the phase on the geometry's uvw, or
    None; gains: device complex128 [nant, nchan] with ant1 / ant2 int32 [nrow] (the tensors
    must outlive the call); psf: invert the point spread function; image_order 1: images in
    [pol, y, x] order; scale: per image polarisation factor of the invert (order 1)."""
    import torch

    c = _lib.HipngColumns()
    c.nchan, c.chan0, c.npol = int(nchan), int(chan0), int(npol)
    c.vis_type = {None: 0, torch.complex128: 0, torch.complex64: 1}[vis_dtype]
    c.weight_type = {None: 0, torch.float64: 0, torch.float32: 1}[weight_dtype]
    c.flag_type = {None: 0, torch.int32: 0, torch.int8: 1}[flag_dtype]
    cm = numpy.ascontiguousarray(numpy.asarray(conv, dtype=complex))
    if cm.ndim != 2 or cm.size > 16:
        raise ValueError("conv must be a complex matrix with at most 16 entries")
    flat = cm.view(numpy.float64).ravel()
    for k in range(flat.size):
        c.conv[k] = float(flat[k])
    if lmn is not None:
        c.shift = 1
        for k in range(3):
            c.lmn[k] = float(lmn[k])
    if gains is not None:
        if gains.dtype != torch.complex128 or not gains.is_contiguous() or gains.dim() != 2 or gains.shape[1] != nchan:
            raise ValueError("gains must be a contiguous complex128 tensor [nant, nchan]")
        for name, a in (("ant1", ant1), ("ant2", ant2)):
            if a is None or a.dtype != torch.int32 or not a.is_contiguous():
                raise ValueError(f"{name} must be a contiguous int32 tensor [nrow]")
        c.nant = int(gains.shape[0])
        c.gains, c.ant1, c.ant2 = gains.data_ptr(), ant1.data_ptr(), ant2.data_ptr()
    c.psf = 1 if psf else 0
    c.image_order = int(image_order)
    c.image_type = {None: 0, torch.float64: 0, torch.float32: 1}[image_dtype]
    sc = [1.0, 1.0, 1.0, 1.0] if scale is None else list(scale) + [1.0] * (4 - len(scale))
    for k in range(4):
        c.scale[k] = float(sc[k])
    return c

