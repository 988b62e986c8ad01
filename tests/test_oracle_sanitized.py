"""The C half of the oracle (oracle/wstack_ref.c: keys, bucket sort, records, per-plane
gridding / degridding, screens) built with AddressSanitizer + UndefinedBehaviorSanitizer
and driven through the CPU port's invert and predict and the bucket sort on a small
observation in a child process: no sanitizer report, and the same results as the plain
build."""

import json
import os
import subprocess
import sys

import numpy
import pytest

from conftest import rel_l2
from oracle import build as obuild
from oracle import sortref, wstack_cpu
from rascil_amd import wgridder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FREQ = numpy.array([1.0e8, 1.17e8])

CHILD = r"""
import json, sys
import numpy
sys.path.insert(0, sys.argv[1])
from oracle import sortref, wstack_cpu
d = numpy.load(sys.argv[2])
plans = json.load(open(sys.argv[3]))
out = {}
for name, plan in plans.items():
    out["dirty_" + name] = wstack_cpu.cpu_ms2dirty(d["uvw"], d["freq"], d["ms"], d["wgt"], plan)
    out["vis_" + name] = wstack_cpu.cpu_dirty2ms(d["uvw"], d["freq"], d["dirty"], d["wgt"], plan)
    sx, sy, sw, _ = wstack_cpu.channel_tables(plan, d["freq"])
    out["keys_" + name], out["perm_" + name] = sortref.bucket_sort(d["uvw"], sx, sy, sw, plan)
numpy.savez(sys.argv[4], **out)
maps = open("/proc/self/maps").read()
print("child done", "libasan" in maps, "liboracle_san" in maps, flush=True)
"""


def _runtime():
    path = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def _plan(uvw, wst):
    ul = numpy.abs(uvw).max(axis=0) * FREQ.max() / 299792458.0
    wmin = numpy.abs(uvw[:, 2]).min() * FREQ.min() / 299792458.0
    return wgridder.plan_params(ul[0], ul[1], wmin, ul[2], 40, 48, 8e-4, 7e-4, 1e-6, wst, "fp64", 2).as_dict()


def test_oracle_c_under_asan_and_ubsan(small_obs, tmp_path):
    asan = _runtime()
    if asan is None:
        pytest.skip("gcc has no AddressSanitizer runtime here")
    uvw = small_obs[0][:400]
    rng = numpy.random.default_rng(11)
    ms = rng.normal(size=(400, 2, 2)) + 1j * rng.normal(size=(400, 2, 2))
    wgt = rng.uniform(0.5, 1.5, (400, 2, 2))
    dirty = rng.normal(size=(2, 40, 48))
    plans = {"wst": _plan(uvw, True), "2d": _plan(uvw, False)}
    numpy.savez(tmp_path / "in.npz", uvw=uvw, freq=FREQ, ms=ms, wgt=wgt, dirty=dirty)
    with open(tmp_path / "plans.json", "w") as f:
        json.dump(plans, f)
    lib = obuild.build(sanitize=True, output=str(tmp_path / "liboracle_san.so"))
    env = dict(os.environ, LD_PRELOAD=asan, ORACLE_LIB=lib, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    script = tmp_path / "child.py"
    script.write_text(CHILD)
    r = subprocess.run([sys.executable, str(script), ROOT, str(tmp_path / "in.npz"), str(tmp_path / "plans.json"),
                        str(tmp_path / "out.npz")], env=env, capture_output=True, text=True, timeout=300)
    report = r.stderr[-4000:]
    assert r.returncode == 0, report
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert "child done True True" in r.stdout, r.stdout  # the sanitizer runtime and build were loaded
    got = numpy.load(tmp_path / "out.npz")
    for name, plan in plans.items():
        ref_d = wstack_cpu.cpu_ms2dirty(uvw, FREQ, ms, wgt, plan)
        ref_v = wstack_cpu.cpu_dirty2ms(uvw, FREQ, dirty, wgt, plan)
        sx, sy, sw, _ = wstack_cpu.channel_tables(plan, FREQ)
        ks, perm = sortref.bucket_sort(uvw, sx, sy, sw, plan)
        assert rel_l2(got["dirty_" + name], ref_d) < 1e-13
        assert rel_l2(got["vis_" + name], ref_v) < 1e-13
        assert numpy.array_equal(got["keys_" + name], ks) and numpy.array_equal(got["perm_" + name], perm)
