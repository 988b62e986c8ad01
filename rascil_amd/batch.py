"""Many independent imaging jobs on one GPU: a plan autotuner and a batch executor.

RASCIL images many snapshots, channels and facets of uneven size through repeated
invert / predict calls (skymodel/skymodel_imaging.py:52-90,
workflows/rsexecute/imaging/imaging_rsexecute.py:72-157, one Dask task per call). Here:

* AutoTuner picks the plan options (oversampling, kernel support, uv tile) per job class
  by timing the candidate plans that meet the job's epsilon on the job itself (geometry
  build + invert + predict), and reuses the choice for every later job of the class
  (image size, visibilities, channels, precision, epsilon, w-stacking; visibilities and
  channels bucketed by powers of 2).
* Batch builds one geometry per job (with the tuned options), captures each job's invert
  + predict into a graph once, and replays the graphs over several streams so that the
  small jobs' kernels overlap on the GPU instead of running one job at a time.
"""

import collections
import math
import time

import numpy

from . import wgridder

# candidate plan options the tuner times (each must still meet epsilon, or it is skipped)
CANDIDATES = ({"sigma": 1.25}, {"sigma": 1.5}, {"sigma": 1.75}, {"sigma": 2.0}, {"sigma": 1.5, "tile": 32},
              {"sigma": 2.0, "tile": 32})


def job_class(nrow, nchan, npix, precision, epsilon, do_wstacking=True):
    """Class key of a job: image size, visibilities in factor-4 buckets, channels in
    factor-2 buckets, precision, epsilon and w-stacking."""
    nvis = max(1, int(nrow) * int(nchan))
    return (int(npix), int(math.log2(nvis)) // 2, int(math.log2(max(1, nchan))), precision, float(epsilon),
            bool(do_wstacking))


class AutoTuner:
    """Plan options per job class, chosen by timing on the GPU (see module docstring).

    tune(uvw, freq, npix, cell, epsilon, ...) returns the plan options (a dict for
    wgridder's plan_opts) of the job's class, timing the candidates on this job the
    first time the class is seen. `table` holds {class: (options, {candidate: ms})}."""

    def __init__(self, candidates=CANDIDATES, repeats=1):
        self.candidates = [dict(c) for c in candidates]
        self.repeats = int(repeats)
        self.table = {}

    def tune(self, uvw, freq, npix, cell, epsilon, precision="mixed", do_wstacking=True, nbatch=1):
        import torch

        freq = numpy.atleast_1d(numpy.asarray(freq, "f8"))
        key = job_class(len(uvw), len(freq), npix, precision, epsilon, do_wstacking)
        if key in self.table:
            return dict(self.table[key][0])
        dev = torch.device("cuda", torch.cuda.current_device())
        cd = torch.complex128 if precision == "fp64" else torch.complex64
        rd = torch.float64 if precision == "fp64" else torch.float32
        u = uvw if hasattr(uvw, "is_cuda") else torch.tensor(numpy.ascontiguousarray(uvw), device=dev)
        ms = torch.ones((u.shape[0], len(freq), nbatch), dtype=cd, device=dev)
        img = torch.zeros((nbatch, npix, npix), dtype=rd, device=dev)
        times = {}
        best = None
        for cand in self.candidates:
            try:
                t_best = None
                for _ in range(self.repeats + 1):  # the first run also builds rocFFT plans
                    torch.cuda.synchronize()
                    t0 = time.time()
                    g = wgridder.Geometry(u, freq, npix, npix, cell, cell, epsilon, do_wstacking, precision,
                                          nbatch=nbatch, plan_opts=cand)
                    g.ms2dirty(ms)
                    g.dirty2ms(img)
                    torch.cuda.synchronize()
                    dt = (time.time() - t0) * 1e3
                    g.close()
                    t_best = dt if t_best is None else min(t_best, dt)
            except ValueError:
                continue  # this candidate cannot reach epsilon (or does not fit)
            times[_cand_name(cand)] = round(t_best, 3)
            if best is None or t_best < best[1]:
                best = (cand, t_best)
        if best is None:
            raise ValueError(f"no candidate plan reaches epsilon {epsilon} for this job")
        self.table[key] = (dict(best[0]), times)
        return dict(best[0])


def _cand_name(c):
    return ",".join(f"{k}={c[k]}" for k in sorted(c))


class Batch:
    """Independent invert + predict jobs captured once and replayed concurrently.

    jobs: list of dicts with uvw [nrow, 3] (metres, ducc0 convention), freq [nchan] Hz,
    npix, cell (rad), and optionally ms [nrow, nchan] / model [npix, npix] (GPU tensors of
    the job's precision; default: ones and zeros) and plan_opts. One geometry per job (plan
    options from `tuner` per job class when given), one graph per job; run() replays every
    graph, job k on stream k % nstreams, and returns per-job latencies and the total time.
    Results stay in self.dirty[k] / self.vis[k]."""

    def __init__(self, jobs, epsilon, precision="mixed", do_wstacking=True, tuner=None, nstreams=4):
        import torch

        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.jobs = jobs
        self.nstreams = max(1, int(nstreams))
        cd = torch.complex128 if precision == "fp64" else torch.complex64
        rd = torch.float64 if precision == "fp64" else torch.float32
        self.geoms, self.graphs, self.ms, self.model, self.dirty, self.vis, self.opts = [], [], [], [], [], [], []
        self.streams = [torch.cuda.Stream() for _ in range(self.nstreams)]
        for j in jobs:
            freq = numpy.atleast_1d(numpy.asarray(j["freq"], "f8"))
            opts = j.get("plan_opts")
            if opts is None and tuner is not None:
                opts = tuner.tune(j["uvw"], freq, j["npix"], j["cell"], epsilon, precision, do_wstacking)
            g = wgridder.Geometry(j["uvw"], freq, j["npix"], j["npix"], j["cell"], j["cell"], epsilon,
                                  do_wstacking, precision, plan_opts=opts)
            n = len(j["uvw"])
            ms = j.get("ms")
            ms = torch.ones((n, len(freq), 1), dtype=cd, device=self.dev) if ms is None else ms.reshape(
                n, len(freq), 1).to(cd).contiguous()
            mod = j.get("model")
            mod = torch.zeros((1, j["npix"], j["npix"]), dtype=rd, device=self.dev) if mod is None else mod.reshape(
                1, j["npix"], j["npix"]).to(rd).contiguous()
            self.geoms.append(g)
            self.opts.append(opts)
            self.ms.append(ms)
            self.model.append(mod)
            self.dirty.append(torch.empty((1, j["npix"], j["npix"]), dtype=rd, device=self.dev))
            self.vis.append(torch.empty((n, len(freq), 1), dtype=cd, device=self.dev))
        torch.cuda.synchronize()
        cap = torch.cuda.Stream()
        with torch.cuda.stream(cap):
            for k, g in enumerate(self.geoms):  # warm-up outside the capture
                g.ms2dirty(self.ms[k], out=self.dirty[k])
                g.dirty2ms(self.model[k], out=self.vis[k])
        cap.synchronize()
        for k, g in enumerate(self.geoms):
            gph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gph, stream=cap):
                g.ms2dirty(self.ms[k], out=self.dirty[k])
                g.dirty2ms(self.model[k], out=self.vis[k])
            self.graphs.append(gph)
        torch.cuda.synchronize()

    def run(self):
        """Replay every job; returns {"seconds": total wall time, "latency_ms": [per job]}: a
        job's latency is the span of its own graph on its stream (from its start to its
        end), "queued_ms" from the start of the batch to its end."""
        import torch

        start = torch.cuda.Event(enable_timing=True)
        begins = [torch.cuda.Event(enable_timing=True) for _ in self.graphs]
        ends = [torch.cuda.Event(enable_timing=True) for _ in self.graphs]
        torch.cuda.synchronize()
        t0 = time.time()
        start.record(torch.cuda.current_stream())
        for s in self.streams:
            s.wait_event(start)
        # longest jobs first, dealt round-robin to the streams
        order = sorted(range(len(self.graphs)), key=lambda k: -self.jobs[k].get("nvis", len(self.jobs[k]["uvw"])))
        for i, k in enumerate(order):
            s = self.streams[i % self.nstreams]
            with torch.cuda.stream(s):
                begins[k].record(s)
                self.graphs[k].replay()
                ends[k].record(s)
        for s in self.streams:
            torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        total = time.time() - t0
        return {"seconds": total, "latency_ms": [b.elapsed_time(e) for b, e in zip(begins, ends)],
                "queued_ms": [start.elapsed_time(e) for e in ends]}

    def close(self):
        for g in self.geoms:
            g.close()
        self.geoms, self.graphs = [], []


def class_table(tuner):
    """The tuner's choices as plain data (for reports)."""
    out = collections.OrderedDict()
    for key, (opts, times) in tuner.table.items():
        out["npix=%d log4vis=%d log2chan=%d %s eps=%g wst=%d" % (key[0], key[1], key[2], key[3], key[4], key[5])] = {
            "chosen": opts, "ms": times}
    return out
