"""hipng_apply_gaintable_noalloc (prepared gain matrices in the library's static per-GPU
array, no device allocation) against hipng_apply_gaintable (library workspace) through the
C-ABI: bit-identical visibilities and weights on two calls in a row, the second with a
smaller input, and a gain table beyond the array's size refused."""

import ctypes

import numpy
import pytest

pytestmark = pytest.mark.gpu


def _case(rng, ntimes, nant, nchan, npol, ngain, gnchan, inverse):
    import torch

    dev = torch.device("cuda", 0)
    nrec = 1 if npol == 1 else 2
    a1, a2 = numpy.triu_indices(nant, 1)
    nbl = len(a1)
    shape = (ntimes, nbl, nchan, npol)
    vis = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    wt = rng.uniform(0.5, 2.0, shape)
    gain = (1 + 0.1 * (rng.normal(size=(ngain, nant, gnchan, nrec, nrec))
                       + 1j * rng.normal(size=(ngain, nant, gnchan, nrec, nrec))))
    if nrec == 2:
        gain[..., 0, 1] *= 0.1
        gain[..., 1, 0] *= 0.1
        gain[0, 1, 0] = [[1, 2], [2, 4]]  # singular: the inverse zeroes those visibilities
    gw = rng.uniform(0.5, 1.0, gain.shape)
    t2g = rng.integers(-1, ngain, ntimes).astype(numpy.int32)
    flags = (rng.uniform(size=shape) < 0.1).astype(numpy.float64)
    tflag = (rng.uniform(size=ntimes) < 0.5).astype(numpy.int8)
    t = lambda x: torch.tensor(numpy.ascontiguousarray(x), device=dev)  # noqa: E731
    return dict(vis=t(vis), wt=t(wt), flags=t(flags), tflag=t(tflag), a1=t(a1.astype(numpy.int32)),
                a2=t(a2.astype(numpy.int32)), gain=t(gain), gw=t(gw), t2g=t(t2g), ntimes=ntimes, nbl=nbl,
                nchan=nchan, npol=npol, ngain=ngain, nant=nant, gnchan=gnchan, nrec=nrec, inverse=int(inverse))


def _call(fn, c, vis, wt):
    import torch

    vp = ctypes.c_void_p
    return fn(vp(vis.data_ptr()), vp(wt.data_ptr()), vp(c["flags"].data_ptr()), vp(c["tflag"].data_ptr()),
              c["ntimes"], c["nbl"], c["nchan"], c["npol"], vp(c["a1"].data_ptr()), vp(c["a2"].data_ptr()),
              vp(c["gain"].data_ptr()), vp(c["gw"].data_ptr()), c["ngain"], c["nant"], c["gnchan"], c["nrec"],
              vp(c["t2g"].data_ptr()), c["inverse"], vp(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("npol,inverse", [(4, False), (4, True), (2, True), (1, False)])
def test_noalloc_gains_equal_workspace_gains_twice(npol, inverse):
    import torch

    from rascil_amd import _lib

    lib = _lib.load()
    rng = numpy.random.default_rng(70 + npol)
    # a large table first, then a smaller one: the second call must not see the first's entries
    for ntimes, nant, nchan, ngain, gnchan in [(24, 40, 8, 6, 8), (9, 12, 3, 2, 1)]:
        c = _case(rng, ntimes, nant, nchan, npol, ngain, gnchan, inverse)
        v_ref, w_ref = c["vis"].clone(), c["wt"].clone()
        v_new, w_new = c["vis"].clone(), c["wt"].clone()
        _lib.check(_call(lib.hipng_apply_gaintable, c, v_ref, w_ref))
        _lib.check(_call(lib.hipng_apply_gaintable_noalloc, c, v_new, w_new))
        torch.cuda.synchronize()
        assert not torch.equal(v_ref, c["vis"])  # the gains did something
        assert torch.equal(v_new, v_ref) and torch.equal(w_new, w_ref)


def test_noalloc_gains_refuse_a_table_beyond_the_static_array():
    from rascil_amd import _lib

    lib = _lib.load()
    rng = numpy.random.default_rng(80)
    c = _case(rng, 2, 4, 1, 4, 1, 1, False)
    c["ngain"], c["nant"], c["gnchan"] = 1, 4, 40000  # 160k entries x 68 B > 8 MiB
    with pytest.raises(ValueError, match="allocation-free"):
        _lib.check(_call(lib.hipng_apply_gaintable_noalloc, c, c["vis"], c["wt"]))
