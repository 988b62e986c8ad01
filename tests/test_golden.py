"""Committed golden fixtures (tests/golden/, written by tests/golden/make_golden.py):
CPU checks that the oracle and the host planning code still reproduce them, GPU checks
that the HIP path reproduces them through the C-ABI."""

import json
import os

import numpy
import pytest

from conftest import rel_l2
from oracle import explicit as ox
from oracle import sortref
from oracle import wstack_cpu
from rascil_amd import _lib, wgridder

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C = 299792458.0


def _ng():
    with numpy.load(os.path.join(GOLD, "ng_explicit.npz")) as z:
        return {k: z[k] for k in z.files}


def _sort():
    with numpy.load(os.path.join(GOLD, "sort_c0.npz")) as z:
        arrs = {k: z[k] for k in z.files}
    with open(os.path.join(GOLD, "sort_c0_plan.json")) as f:
        plan = json.load(f)
    return arrs, plan


def _plan_struct(d):
    """hipng_plan from the fixture's fields (diagnostic fields added later stay 0)."""
    p = _lib.HipngPlan()
    for name, _ in p._fields_:
        setattr(p, name, d.get(name, 0))
    return p


@pytest.mark.parametrize("tag,wst", [("wst", True), ("2d", False)])
def test_oracle_reproduces_ng_fixture(tag, wst):
    g = _ng()
    (nx, ny), (px, py) = g["npix"], g["pixsize"]
    d = ox.explicit_ms2dirty(g["uvw"], g["freq"], g["ms"], g["wgt"], nx, ny, px, py, wst)
    v = ox.explicit_dirty2ms(g["uvw"], g["freq"], g["model"], g["wgt"], px, py, wst)
    numpy.testing.assert_allclose(d, g["dirty_" + tag], rtol=0, atol=1e-12 * numpy.abs(d).max())
    numpy.testing.assert_allclose(v, g["vis_" + tag], rtol=0, atol=1e-12 * numpy.abs(v).max())


@pytest.mark.parametrize("tag", ["wst", "2d"])
def test_ng_fixture_adjoint_pair(tag):
    """<model, dirty(ms)> = Re sum ms conj(vis(model)) for the stored weighted pair."""
    g = _ng()
    lhs = float(numpy.sum(g["model"] * g["dirty_" + tag]))
    rhs = float(numpy.sum(g["ms"] * numpy.conj(g["vis_" + tag])).real)
    assert lhs == pytest.approx(rhs, rel=1e-11)


def test_ng_fixture_wstacking_differs_from_2d():
    """The fixture's field of view is wide enough that the w term and 1/n matter."""
    g = _ng()
    assert rel_l2(g["dirty_2d"], g["dirty_wst"]) > 1e-3
    assert rel_l2(g["vis_2d"], g["vis_wst"]) > 1e-3


@pytest.mark.parametrize("tag,wst", [("wst", True), ("2d", False)])
def test_cpu_port_reproduces_ng_fixture(tag, wst):
    g = _ng()
    (nx, ny), (px, py) = g["npix"], g["pixsize"]
    ul = numpy.abs(g["uvw"]).max(axis=0) * g["freq"].max() / C
    wmin = numpy.abs(g["uvw"][:, 2]).min() * g["freq"].min() / C
    plan = wgridder.plan_params(ul[0], ul[1], wmin, ul[2], int(nx), int(ny), px, py, 1e-9, wst,
                                "fp64").as_dict()
    d = wstack_cpu.cpu_ms2dirty(g["uvw"], g["freq"], g["ms"], g["wgt"], plan)[0]
    v = wstack_cpu.cpu_dirty2ms(g["uvw"], g["freq"], g["model"], g["wgt"], plan)[..., 0]
    assert rel_l2(d, g["dirty_" + tag]) < 1e-9
    assert rel_l2(v, g["vis_" + tag]) < 1e-9


def test_plan_and_scales_reproduce_sort_fixture():
    a, plan = _sort()
    ul = numpy.abs(a["uvw"]).max(axis=0) * a["freq"].max() / C
    wmin = numpy.abs(a["uvw"][:, 2]).min() * a["freq"].min() / C
    p = wgridder.plan_params(ul[0], ul[1], wmin, ul[2], 128, 128, 6e-4, 6e-4, 1e-5, True, "fp64")
    got = p.as_dict()
    # every field the fixture holds (fields added to the plan since are diagnostics); the
    # modelled error is no part of the sort contract and now also carries the transforms'
    # roundoff floor (1e-12 here)
    assert {k: got[k] for k in plan if k != "est_error"} == {k: v for k, v in plan.items() if k != "est_error"}
    assert got["est_error"] == pytest.approx(plan["est_error"], rel=1e-5) and got["est_error"] <= 1e-5
    sx, sy, sw = wgridder.channel_scales(p, a["freq"])
    assert numpy.array_equal(sx, a["sx"]) and numpy.array_equal(sy, a["sy"])
    assert numpy.array_equal(sw, a["sw"])


def test_sort_oracle_reproduces_fixture():
    a, plan = _sort()
    keys, perm = sortref.bucket_sort(a["uvw"], a["sx"], a["sy"], a["sw"], plan)
    assert numpy.array_equal(keys, a["keys"])
    assert numpy.array_equal(perm, a["perm"])
    assert numpy.all(numpy.diff(a["keys"].astype(numpy.int64)) >= 0)
    assert numpy.array_equal(numpy.sort(a["perm"]), numpy.arange(a["perm"].size))


@pytest.mark.gpu
def test_gpu_sort_reproduces_fixture():
    a, plan = _sort()
    keys, perm = wgridder.bucket_sort(a["uvw"], a["freq"], _plan_struct(plan))
    assert numpy.array_equal(keys.astype(numpy.uint32), a["keys"])
    assert numpy.array_equal(perm.astype(numpy.uint32), a["perm"])


@pytest.mark.gpu
@pytest.mark.parametrize("prec,eps,tol", [("fp64", 1e-12, 1e-11), ("fp64", 1e-7, 1e-7),
                                          ("mixed", 1e-5, 1e-5), ("fp32", 1e-4, 1e-4)])
@pytest.mark.parametrize("tag,wst", [("wst", True), ("2d", False)])
def test_gpu_reproduces_ng_fixture(prec, eps, tol, tag, wst):
    g = _ng()
    (nx, ny), (px, py) = g["npix"], g["pixsize"]
    d = wgridder.ms2dirty(g["uvw"], g["freq"], g["ms"], g["wgt"], int(nx), int(ny), px, py, 0, 0, eps,
                          wst, precision=prec)
    v = wgridder.dirty2ms(g["uvw"], g["freq"], g["model"], g["wgt"], px, py, 0, 0, eps, wst,
                          precision=prec)
    assert rel_l2(numpy.asarray(d, numpy.float64), g["dirty_" + tag]) < tol
    assert rel_l2(numpy.asarray(v, numpy.complex128), g["vis_" + tag]) < tol
