"""ORACLE — test infrastructure only: ctypes wrapper of oracle/liboracle.so (bucket sort)."""

import ctypes
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OraclePlan(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("gx", "gy", "W", "tile", "ntx", "nty", "nplanes", "wstack", "f32coords")] + \
               [("w0", ctypes.c_double), ("dw", ctypes.c_double)]


def lib():
    """The oracle's C library: oracle/liboracle.so (built on first use), or the file named
    by ORACLE_LIB (the sanitizer build of the tests)."""
    global _LIB
    if _LIB is None:
        path = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build

            build.build()
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def bucket_sort(uvw, sx, sy, sw, plan):
    """CPU keys + stable sort by (bucket key, in-tile origin digit). `plan` holds the
    hipng_plan fields (dict-like). Returns the sorted bucket keys and the permutation."""
    uvw = numpy.ascontiguousarray(uvw, "f8")
    sx, sy, sw = (numpy.ascontiguousarray(a, "f8") for a in (sx, sy, sw))
    nrow, nchan = uvw.shape[0], sx.shape[0]
    op = OraclePlan(plan["gx"], plan["gy"], plan["W"], plan["tile"], plan["ntx"], plan["nty"],
                    plan["nplanes"], plan["wstacking"], 1 if plan["precision"] == 2 else 0,
                    plan["w0"], plan["dw"])
    n = nrow * nchan
    keys = numpy.empty(n, numpy.uint32)
    digits = numpy.empty(n, numpy.uint32)
    L = lib()
    L.oracle_keys(_p(uvw), ctypes.c_int64(nrow), ctypes.c_int32(nchan), _p(sx), _p(sy), _p(sw),
                  ctypes.byref(op), _p(keys), _p(digits))
    nkeys = plan["nplanes"] * plan["ntx"] * plan["nty"]
    ks = numpy.empty(n, numpy.uint32)
    perm = numpy.empty(n, numpy.uint32)
    rc = L.oracle_bucket_sort(_p(keys), _p(digits), ctypes.c_int64(n), ctypes.c_uint64(nkeys),
                              ctypes.c_uint64(plan["tile"] * plan["tile"]), _p(ks), _p(perm))
    if rc != 0:
        raise MemoryError("oracle counting sort allocation failed")
    return ks, perm
