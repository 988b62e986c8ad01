"""Writes the committed golden fixtures of this directory (run from the repo root:
``python tests/golden/make_golden.py``).

* ng_explicit.npz: a 400-row subsample of the 100-station / 1 km / dec -45 deg geometry
  at 2 channels (100, 131 MHz), seeded random visibilities and weights, a 48 x 40 image
  (pixel sizes 8e-4 / 9e-4 rad). Expected outputs are the FP64 brute-force sums that
  ducc0.wgridder.ms2dirty / dirty2ms approximate (called from
  rascil/processing_components/imaging/ng.py:91-120, 232-275), with and without
  w-stacking.
* sort_c0.npz + sort_c0_plan.json: 2,000 rows of the same geometry at 3 channels, the plan
  the library chooses for a 128^2, 6e-4 rad image at epsilon 1e-5 with w-stacking, and the
  bucket keys and stable permutation the CPU sort oracle computes for it.

The fixtures are data (inputs and expected outputs); the oracle is only run here to make
them.
"""

import json
import os
import sys

import numpy

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import explicit as ox  # noqa: E402
from oracle import sortref  # noqa: E402
from rascil_amd import simulation, wgridder  # noqa: E402

C = 299792458.0


def _rows(n, seed):
    bvis = simulation.config0(seed=0)
    uvw = bvis.uvw.reshape(-1, 3)
    rng = numpy.random.default_rng(seed)
    return uvw[numpy.sort(rng.choice(uvw.shape[0], n, replace=False))].copy()


def make_ng(path):
    uvw = _rows(400, 21)
    freq = numpy.array([1.0e8, 1.31e8])
    rng = numpy.random.default_rng(22)
    ms = rng.normal(size=(400, 2)) + 1j * rng.normal(size=(400, 2))
    wgt = rng.uniform(0.2, 1.5, (400, 2))
    nx, ny, px, py = 48, 40, 8e-4, 9e-4
    model = rng.normal(size=(nx, ny))
    out = dict(uvw=uvw, freq=freq, ms=ms, wgt=wgt, npix=numpy.array([nx, ny]),
               pixsize=numpy.array([px, py]), model=model)
    for tag, wst in (("wst", True), ("2d", False)):
        out["dirty_" + tag] = ox.explicit_ms2dirty(uvw, freq, ms, wgt, nx, ny, px, py, wst)
        out["vis_" + tag] = ox.explicit_dirty2ms(uvw, freq, model, wgt, px, py, wst)
    numpy.savez_compressed(path, **out)


def make_sort(path, plan_path):
    uvw = _rows(2000, 23)
    freq = numpy.array([1.0e8, 1.07e8, 1.31e8])
    ul = numpy.abs(uvw).max(axis=0) * freq.max() / C
    wmin = numpy.abs(uvw[:, 2]).min() * freq.min() / C
    plan = wgridder.plan_params(ul[0], ul[1], wmin, ul[2], 128, 128, 6e-4, 6e-4, 1e-5, True, "fp64")
    sx, sy, sw = wgridder.channel_scales(plan, freq)
    keys, perm = sortref.bucket_sort(uvw, sx, sy, sw, plan.as_dict())
    numpy.savez_compressed(path, uvw=uvw, freq=freq, sx=sx, sy=sy, sw=sw, keys=keys, perm=perm)
    with open(plan_path, "w") as f:
        json.dump(plan.as_dict(), f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    make_ng(os.path.join(HERE, "ng_explicit.npz"))
    make_sort(os.path.join(HERE, "sort_c0.npz"), os.path.join(HERE, "sort_c0_plan.json"))
    print("fixtures written to", HERE)
