"""Loader for the hipng C-ABI shared library (``rascil_amd/_hipng.so``).

The library is the only compute path of this package: when it is missing or fails to
load, every entry point raises instead of falling back to anything else.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "_hipng.so")

HIPNG_FP64, HIPNG_MIXED, HIPNG_FP32 = 0, 1, 2
HIPNG_OK, HIPNG_EINVAL, HIPNG_ERUNTIME = 0, -1, -2

PRECISIONS = {"fp64": HIPNG_FP64, "mixed": HIPNG_MIXED, "fp32": HIPNG_FP32}

# Every function declared in include/hipng.h
EXPORTS = (
    "hipng_version",
    "hipng_last_error",
    "hipng_plan_params",
    "hipng_channel_scales",
    "hipng_uvw_extent",
    "hipng_bucket_sort",
    "hipng_ms2dirty",
    "hipng_dirty2ms",
    "hipng_geometry_create",
    "hipng_geometry_ms2dirty",
    "hipng_geometry_dirty2ms",
    "hipng_geometry_destroy",
    "hipng_geometry_stats",
    "hipng_geometry_invert_columns",
    "hipng_geometry_predict_columns",
    "hipng_weight_sums",
    "hipng_gain_gradient",
    "hipng_dft",
    "hipng_idft",
    "hipng_weight_visibility",
    "hipng_apply_gaintable",
    "hipng_apply_gaintable_noalloc",
    "hipng_vis_prologue",
    "hipng_vis_epilogue",
    "hipng_profile_enable",
    "hipng_profile_query",
    "hipng_release",
)


class HipngPlan(ctypes.Structure):
    """Mirror of ``hipng_plan`` in include/hipng.h."""

    _fields_ = [
        ("precision", ctypes.c_int32),
        ("wstacking", ctypes.c_int32),
        ("W", ctypes.c_int32),
        ("tile", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("gx", ctypes.c_int32),
        ("gy", ctypes.c_int32),
        ("ntx", ctypes.c_int32),
        ("nty", ctypes.c_int32),
        ("nplanes", ctypes.c_int32),
        ("keybits", ctypes.c_int32),
        ("beta", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("pixsize_x", ctypes.c_double),
        ("pixsize_y", ctypes.c_double),
        ("w0", ctypes.c_double),
        ("dw", ctypes.c_double),
        ("nshift", ctypes.c_double),
        ("umax", ctypes.c_double),
        ("vmax", ctypes.c_double),
        ("wmin", ctypes.c_double),
        ("wmax", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("est_error", ctypes.c_double),
        ("dense_min", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("dense_items", ctypes.c_int64),
        ("occupied_tiles", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class HipngPlanOpts(ctypes.Structure):
    """Mirror of ``hipng_plan_opts`` in include/hipng.h: plan choices pinned by the caller
    (0 = automatic)."""

    _fields_ = [
        ("sigma", ctypes.c_double),
        ("W", ctypes.c_int32),
        ("tile", ctypes.c_int32),
        ("dense_min", ctypes.c_int32),
        ("chunk_grid", ctypes.c_int32),
        ("chunk_degrid", ctypes.c_int32),
        ("sort_pack", ctypes.c_int32),
        ("use_extent", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("umax", ctypes.c_double),
        ("vmax", ctypes.c_double),
        ("wmin", ctypes.c_double),
        ("wmax", ctypes.c_double),
    ]

    FIELDS = ("sigma", "W", "tile", "dense_min", "chunk_grid", "chunk_degrid", "sort_pack", "extent")


HIPNG_COMPLEX128, HIPNG_COMPLEX64 = 0, 1
HIPNG_FLOAT64, HIPNG_FLOAT32 = 0, 1
HIPNG_INT32, HIPNG_INT8 = 0, 1


class HipngColumns(ctypes.Structure):
    """Mirror of ``hipng_columns`` in include/hipng.h."""

    _fields_ = [
        ("nchan", ctypes.c_int32),
        ("chan0", ctypes.c_int32),
        ("npol", ctypes.c_int32),
        ("vis_type", ctypes.c_int32),
        ("weight_type", ctypes.c_int32),
        ("flag_type", ctypes.c_int32),
        ("shift", ctypes.c_int32),
        ("nant", ctypes.c_int32),
        ("lmn", ctypes.c_double * 3),
        ("conv", ctypes.c_double * 32),
        ("psf", ctypes.c_int32),
        ("image_order", ctypes.c_int32),
        ("image_type", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("scale", ctypes.c_double * 4),
        ("gains", ctypes.c_void_p),
        ("ant1", ctypes.c_void_p),
        ("ant2", ctypes.c_void_p),
    ]


def plan_opts(opts):
    """ctypes pointer to a HipngPlanOpts built from a dict (or None for all-automatic).

    Keys: sigma, W, tile, dense_min (-1 turns the dense-run path off), chunk_grid,
    chunk_degrid, sort_pack (False recomputes the bucket keys), extent (umax, vmax, wmin,
    wmax) in wavelengths that the plan covers at least."""
    if opts is None:
        return None
    if isinstance(opts, HipngPlanOpts):
        return ctypes.byref(opts)
    unknown = set(opts) - set(HipngPlanOpts.FIELDS)
    if unknown:
        raise ValueError(f"unknown plan options {sorted(unknown)}; use {list(HipngPlanOpts.FIELDS)}")
    o = HipngPlanOpts()
    for k in ("sigma",):
        if opts.get(k) is not None:
            o.sigma = float(opts[k])
    for k in ("W", "tile", "dense_min", "chunk_grid", "chunk_degrid"):
        if opts.get(k) is not None:
            setattr(o, k, int(opts[k]))
    if opts.get("sort_pack") is not None and not opts["sort_pack"]:
        o.sort_pack = -1
    if opts.get("extent") is not None:
        o.use_extent = 1
        o.umax, o.vmax, o.wmin, o.wmax = (float(x) for x in opts["extent"])
    return ctypes.byref(o)


_lib = None
_lock = threading.Lock()


class HipngLibraryMissing(RuntimeError):
    """The compiled hipng library is not available."""


def load():
    """Load (once) and return the ctypes handle of the hipng library."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIBPATH):
            raise HipngLibraryMissing(
                f"hipng library not built: {LIBPATH} is missing "
                "(run rascil_amd.build.build() or __graft_entry__.build())"
            )
        try:
            # torch ships its own HIP, HSA and rocFFT runtimes; loaded first, they are the ones
            # the library binds to. Loaded after the library, they would be a second copy in
            # the process, whose device initialisation fails ("no ROCm-capable device").
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = ctypes.CDLL(LIBPATH)
        c_i32, c_i64, c_dbl, c_vp = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
        pd = ctypes.POINTER(ctypes.c_double)
        pplan = ctypes.POINTER(HipngPlan)
        popts = ctypes.POINTER(HipngPlanOpts)
        lib.hipng_version.restype = c_i32
        lib.hipng_last_error.restype = ctypes.c_char_p
        lib.hipng_plan_params.argtypes = [c_dbl, c_dbl, c_dbl, c_dbl, c_i32, c_i32, c_dbl, c_dbl,
                                          c_dbl, c_i32, c_i32, c_i32, c_i64, popts, pplan]
        lib.hipng_plan_params.restype = c_i32
        lib.hipng_channel_scales.argtypes = [pplan, pd, c_i32, pd, pd, pd]
        lib.hipng_channel_scales.restype = c_i32
        lib.hipng_uvw_extent.argtypes = [c_vp, c_i64, pd, c_i32, pd, c_vp]
        lib.hipng_uvw_extent.restype = c_i32
        lib.hipng_bucket_sort.argtypes = [c_vp, pd, c_i64, c_i32, pplan, popts, c_vp, c_vp, c_vp]
        lib.hipng_bucket_sort.restype = c_i32
        for name in ("hipng_ms2dirty", "hipng_dirty2ms"):
            fn = getattr(lib, name)
            fn.argtypes = [c_vp, pd, c_i64, c_i32, c_vp, c_vp, c_i32, c_i32, c_i32, c_dbl, c_dbl,
                           c_dbl, c_i32, c_i32, c_vp, c_vp, popts, pplan]
            fn.restype = c_i32
        lib.hipng_geometry_create.argtypes = [c_vp, pd, c_i64, c_i32, c_i32, c_i32, c_i32, c_dbl,
                                              c_dbl, c_dbl, c_i32, c_i32, c_vp, popts,
                                              ctypes.POINTER(c_vp), pplan]
        lib.hipng_geometry_create.restype = c_i32
        lib.hipng_geometry_ms2dirty.argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp]
        lib.hipng_geometry_ms2dirty.restype = c_i32
        lib.hipng_geometry_dirty2ms.argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp]
        lib.hipng_geometry_dirty2ms.restype = c_i32
        lib.hipng_geometry_destroy.argtypes = [c_vp]
        lib.hipng_geometry_destroy.restype = c_i32
        lib.hipng_geometry_stats.argtypes = [c_vp, ctypes.POINTER(c_i64), c_i32]
        lib.hipng_geometry_stats.restype = c_i32
        pcols = ctypes.POINTER(HipngColumns)
        lib.hipng_geometry_invert_columns.argtypes = [c_vp, pcols, c_vp, c_vp, c_vp, c_vp, c_vp]
        lib.hipng_geometry_invert_columns.restype = c_i32
        lib.hipng_geometry_predict_columns.argtypes = [c_vp, pcols, c_vp, c_vp, c_vp, c_vp]
        lib.hipng_geometry_predict_columns.restype = c_i32
        lib.hipng_weight_sums.argtypes = [c_vp, c_vp, c_i64, c_i32, c_i32, c_i32, c_i32, c_vp, c_vp]
        lib.hipng_weight_sums.restype = c_i32
        lib.hipng_gain_gradient.argtypes = [c_vp, c_vp, c_i32, c_vp, c_vp, c_vp, c_i64, c_i32, c_i32, c_i32, c_vp,
                                            c_vp]
        lib.hipng_gain_gradient.restype = c_i32
        lib.hipng_dft.argtypes = [c_vp, c_vp, c_vp, c_i32, c_i64, c_i32, c_i32, c_vp, c_vp]
        lib.hipng_dft.restype = c_i32
        lib.hipng_idft.argtypes = [c_vp, c_vp, c_vp, c_vp, c_vp, c_i32, c_i64, c_i32, c_i32, c_vp, c_vp, c_vp]
        lib.hipng_idft.restype = c_i32
        lib.hipng_release.restype = c_i32
        lib.hipng_weight_visibility.argtypes = [c_vp, pd, c_i64, c_i64, c_i32, c_i32,
                                                ctypes.POINTER(ctypes.c_int32), c_i32, c_i32, c_i32,
                                                c_dbl, c_dbl, c_i32, c_dbl, c_vp, c_vp, c_vp]
        lib.hipng_weight_visibility.restype = c_i32
        lib.hipng_apply_gaintable.argtypes = [c_vp, c_vp, c_vp, c_vp, c_i64, c_i32, c_i32, c_i32, c_vp, c_vp,
                                              c_vp, c_vp, c_i32, c_i32, c_i32, c_i32, c_vp, c_i32, c_vp]
        lib.hipng_apply_gaintable.restype = c_i32
        lib.hipng_apply_gaintable_noalloc.argtypes = lib.hipng_apply_gaintable.argtypes
        lib.hipng_apply_gaintable_noalloc.restype = c_i32
        lib.hipng_vis_prologue.argtypes = [c_vp, c_vp, c_vp, c_vp, pd, c_i64, c_i32, c_i32, c_i32, pd, pd, c_i32,
                                           c_vp, c_vp, c_vp, c_vp]
        lib.hipng_vis_prologue.restype = c_i32
        lib.hipng_vis_epilogue.argtypes = [c_vp, c_vp, pd, c_i64, c_i32, c_i32, c_i32, pd, pd, c_i32, c_vp, c_vp]
        lib.hipng_vis_epilogue.restype = c_i32
        lib.hipng_profile_enable.argtypes = [c_i32]
        lib.hipng_profile_enable.restype = c_i32
        lib.hipng_profile_query.argtypes = [c_i32, ctypes.c_char_p, c_i32, pd,
                                            ctypes.POINTER(ctypes.c_int64)]
        lib.hipng_profile_query.restype = c_i32
        _lib = lib
        return lib


def check(code):
    """Turn a hipng return code into the exception ducc0 would raise."""
    if code == HIPNG_OK:
        return
    msg = load().hipng_last_error().decode(errors="replace")
    if code == HIPNG_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def dptr(x):
    """Address of a numpy array's buffer as ctypes double pointer."""
    return x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def profile_enable(on=True):
    """Start (reset) or stop per-kernel HIP-event timing inside the library."""
    check(load().hipng_profile_enable(1 if on else 0))


def profile_read():
    """{kernel class: (total_ms, launches)} accumulated since profile_enable(True)."""
    lib = load()
    out = {}
    i = 0
    name = ctypes.create_string_buffer(64)
    tot = ctypes.c_double()
    cnt = ctypes.c_int64()
    while lib.hipng_profile_query(i, name, 64, ctypes.byref(tot), ctypes.byref(cnt)) == HIPNG_OK:
        out[name.value.decode()] = (tot.value, cnt.value)
        i += 1
    return out
