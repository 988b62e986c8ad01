#!/usr/bin/env python3
"""Throughput of one invert (ms2dirty) + one predict (dirty2ms) at the production shape.

Workload (default "c1"): 512 stations (70 % Gaussian core sigma 500 m, 30 % log-uniform
1-40 km), 128 hour angles over 4 h at dec -30 deg, 64 channels 100-164 MHz, 4
polarisations -> 16,744,448 rows x 64 x 4 = 4.29e9 complex64 visibilities, image
8192^2 at cell 1.1e-5 rad, epsilon 1e-5, w-stacking. Synthetic data (seeded random
visibilities, unit weights, random model image) generated directly in HBM; inputs are
resident before the timed region. One step = hipng ms2dirty of all 4 polarisations +
hipng dirty2ms of a 4-polarisation model image.

Prints one JSON line with value = Mvis/s (nvis / seconds per step), the roofline of the
dominant kernel (algorithmic bytes / average launch time from HIP events recorded by the
library on its launch stream during the timed steps) and the CPU baseline (the oracle's
CPU port of the same algorithm on a bounded sample of the same workload).
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Mvis/s invert+predict, 8192² w-stacked, ε≤1e-5 vs FP64 DFT; FP32 vs FP64 mode"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c1", choices=["c1", "c0"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp32", "fp64"])
    ap.add_argument("--nchan", type=int, default=None, help="override channel count (not the bench line)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-accuracy", action="store_true", help="skip the accuracy probe")
    ap.add_argument("--sigma", type=float, default=None,
                    help="force the grid oversampling factor (cost-model calibration, not the bench line)")
    ap.add_argument("--study", default=None, choices=["precision"],
                    help="run a study instead of the bench line (prints one JSON line)")
    ap.add_argument("--fp64-steps", type=int, default=0,
                    help="also time the FP64 plain mode for this many steps")
    return ap.parse_args()


def workload(cfg, nchan_override):
    from rascil_amd import simulation

    if cfg == "c1":
        uvw, freq, _ = simulation.config1_geometry(nchan=nchan_override or 64)
        return dict(name="c1", uvw=uvw, freq=freq, npol=4, npix=8192, cell=1.1e-5, eps=1e-5,
                    desc=("512 stations (70% Gaussian core sigma 500 m + 30% log-uniform 1-40 km), "
                          f"128 HA over 4 h, dec -30, {len(freq)} ch 100-164 MHz, 4 pol; 8192^2, "
                          "cell 1.1e-5 rad, eps 1e-5, w-stacking"))
    bvis = simulation.config0()
    return dict(name="c0", uvw=bvis.uvw.reshape(-1, 3), freq=bvis.frequency, npol=1, npix=512,
                cell=5e-4, eps=1e-7,
                desc="100 stations 1 km disc, 30 HA, dec -45, 1 ch 100 MHz, Stokes I; 512^2, cell 5e-4, eps 1e-7")


def algorithmic_bytes(name, plan, nitems, nb, csz):
    """HBM bytes one launch of kernel class `name` must move at minimum."""
    G2 = plan["gx"] * plan["gy"]
    N2 = plan["nx"] * plan["ny"]
    P = plan["nplanes"]
    ww = plan["W"] if plan["wstacking"] else 1
    rsz = 4 if csz == 8 else 8
    items_pp = nitems * ww / P
    if name == "grid":  # records of the plane's visibilities in, plane out
        return items_pp * (4 + 3 * rsz + nb * csz) + nb * G2 * csz
    if name == "degrid":  # plane in, records in, sorted accumulators read+write
        return nb * G2 * csz + items_pp * (4 + 3 * rsz + 2 * nb * csz)
    if name.startswith("fft"):  # one in-place read + write of the plane
        return 2 * nb * G2 * csz
    if name == "screen_acc":  # plane pixels in, FP64 accumulator read + write
        return nb * N2 * (csz + 16)
    if name == "screen_fill":  # corrected image in (FP64), plane out
        return nb * N2 * 8 + nb * G2 * csz
    return None


def cpu_plans(plan_i, npol, nitems):
    """Other oversampling choices for the CPU leg (it reports the fastest plan it tried)."""
    from rascil_amd import wgridder

    out = []
    old = os.environ.get("HIPNG_SIGMA")
    try:
        for sg in (1.25, 2.0):
            if abs(sg - plan_i["sigma"]) < 1e-9:
                continue
            os.environ["HIPNG_SIGMA"] = str(sg)
            p = wgridder.plan_params(plan_i["umax"], plan_i["vmax"], plan_i["wmin"], plan_i["wmax"],
                                     plan_i["nx"], plan_i["ny"], plan_i["pixsize_x"], plan_i["pixsize_y"],
                                     plan_i["epsilon"], True, "fp64", npol, nitems).as_dict()
            out.append((p, p))
    finally:
        if old is None:
            os.environ.pop("HIPNG_SIGMA", None)
        else:
            os.environ["HIPNG_SIGMA"] = old
    return out


def accuracy_probe(wl, prec, nsrc=20, frac=0.01, seed=7):
    """Max / RMS relative error of predict and invert at the workload's plan, on a row
    subsample, against FP64 direct sums: predict vs the FP64 HIP DFT of pixel-centred
    sources (flux / n, the w-stacking convention); invert at the source pixels vs the FP64
    adjoint sum evaluated with torch on the GPU."""
    import numpy
    import torch

    from rascil_amd import wgridder
    from rascil_amd.dft import dft_kernel

    rng = numpy.random.default_rng(seed)
    uvw = wl["uvw"]
    rows = numpy.sort(rng.choice(uvw.shape[0], max(1, int(uvw.shape[0] * frac)), replace=False))
    su = numpy.ascontiguousarray(uvw[rows])
    freq, N, cell, eps = wl["freq"], wl["npix"], wl["cell"], wl["eps"]
    rdt = numpy.float64 if prec == "fp64" else numpy.float32
    model = numpy.zeros((N, N), rdt)
    ix = rng.integers(N // 8, N - N // 8, nsrc)
    iy = rng.integers(N // 8, N - N // 8, nsrc)
    flux = rng.uniform(0.1, 1.0, nsrc)
    model[ix, iy] = flux
    # ducc0 convention (the caller negates u and w): x = (ix - N/2) cell -> l = -x, m = y
    x = (ix - N // 2) * cell
    y = (iy - N // 2) * cell
    n = numpy.sqrt(1 - x * x - y * y)
    ruvw = su * numpy.array([-1.0, 1.0, -1.0])
    uvl = (ruvw[:, None, :] * (numpy.asarray(freq) / 299792458.0)[None, :, None])[None]
    dc = numpy.stack([-x, y, n - 1], axis=1)
    ref = dft_kernel(dc, (flux / n)[:, None, None] * numpy.ones((1, len(freq), 1)), uvl)[0, :, :, 0]
    t = time.time()
    vis = wgridder.dirty2ms(su, freq, model, None, cell, cell, 0, 0, eps, True, precision=prec)
    err = numpy.abs(vis - ref)
    out = {"rows": len(rows), "nvis": int(ref.size), "sources": nsrc,
           "predict": {"max_rel": float(err.max() / numpy.abs(ref).max()),
                       "rms_rel": float(numpy.sqrt((err ** 2).mean() / (numpy.abs(ref) ** 2).mean()))}}
    # invert the exact visibilities; compare the source pixels with the FP64 adjoint sum
    dirty = wgridder.ms2dirty(su, freq, ref.astype(vis.dtype), None, N, N, cell, cell, 0, 0, eps, True,
                              precision=prec)
    dev = torch.device("cuda", 0)
    U = torch.tensor(uvl[0].reshape(-1, 3), device=dev)
    V = torch.tensor(ref.reshape(-1), device=dev)
    exact = []
    for k in range(nsrc):
        ph = U[:, 0] * (-x[k]) + U[:, 1] * y[k] + U[:, 2] * (n[k] - 1)  # +phase of the DFT above
        exact.append(float((V * torch.exp(2j * numpy.pi * ph.to(torch.complex128))).real.sum()) / n[k])
    exact = numpy.array(exact)
    got = dirty[ix, iy].astype(numpy.float64)
    e2 = numpy.abs(got - exact)
    out["invert"] = {"max_rel": float(e2.max() / numpy.abs(exact).max()),
                     "rms_rel": float(numpy.sqrt((e2 ** 2).mean() / (exact ** 2).mean()))}
    out["seconds"] = round(time.time() - t, 1)
    return out


def precision_study(seed=3):
    """Three precision modes x epsilon on the production geometry cut to 8 channels and one
    polarisation (134M visibilities), 200 power-law point sources up to 5 deg off-axis at
    pixel centres of an 8192^2, 2.2e-5 rad image. Errors against FP64 direct sums: predicted
    visibilities (all of them) vs the FP64 HIP DFT of flux / n; the dirty image at the
    source pixels vs the FP64 adjoint sum."""
    import numpy
    import torch

    from rascil_amd import simulation, wgridder
    from rascil_amd.dft import dft_kernel

    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(seed)
    uvw, freq, _ = simulation.config1_geometry(nchan=8)
    N, cell, nsrc = 8192, 2.2e-5, 200
    r = numpy.radians(5.0) * numpy.sqrt(rng.uniform(0, 1, nsrc))
    phi = rng.uniform(0, 2 * numpy.pi, nsrc)
    ix = numpy.clip(N // 2 + numpy.round(r * numpy.cos(phi) / cell).astype(int), 0, N - 1)
    iy = numpy.clip(N // 2 + numpy.round(r * numpy.sin(phi) / cell).astype(int), 0, N - 1)
    flux = simulation.powerlaw_fluxes(nsrc, -1.6, 1e-3, 1.0, rng)
    x, y = (ix - N // 2) * cell, (iy - N // 2) * cell
    n = numpy.sqrt(1 - x * x - y * y)
    model = numpy.zeros((N, N))
    numpy.add.at(model, (ix, iy), flux)
    uvl = ((uvw * numpy.array([-1.0, 1.0, -1.0]))[:, None, :] * (freq / 299792458.0)[None, :, None])[None]
    dc = numpy.stack([-x, y, n - 1], axis=1)
    t = time.time()
    vtrue = dft_kernel(dc, (flux / n)[:, None, None] * numpy.ones((1, len(freq), 1)), uvl)[0, :, :, 0]
    print(f"[study] FP64 DFT of {nsrc} sources over {vtrue.size} vis: {time.time() - t:.1f} s",
          file=sys.stderr, flush=True)
    U = torch.tensor(uvl[0].reshape(-1, 3), device=dev)
    V = torch.tensor(vtrue.reshape(-1), device=dev)
    probe = list(range(0, nsrc, 4))
    exact = []
    for k in probe:
        ph = U[:, 0] * (-x[k]) + U[:, 1] * y[k] + U[:, 2] * (n[k] - 1)
        exact.append(float((V * torch.exp(2j * numpy.pi * ph.to(torch.complex128))).real.sum()) / n[k])
    exact = numpy.array(exact)
    del U
    uvw_d = torch.tensor(uvw, device=dev)
    rows = []
    for mode in ("fp64", "mixed", "fp32"):
        for eps in (1e-3, 1e-4, 1e-5, 1e-6):
            cd, rd = (torch.complex128, torch.float64) if mode == "fp64" else (torch.complex64, torch.float32)
            ms = V.to(cd).reshape(len(uvw), len(freq))
            img = torch.tensor(model, dtype=rd, device=dev)
            plan = {}
            torch.cuda.synchronize()
            t0 = time.time()
            dirty = wgridder.ms2dirty(uvw_d, freq, ms, None, N, N, cell, cell, 0, 0, eps, True,
                                      precision=mode, plan_out=plan)
            torch.cuda.synchronize()
            t1 = time.time()
            vis = wgridder.dirty2ms(uvw_d, freq, img, None, cell, cell, 0, 0, eps, True, precision=mode)
            torch.cuda.synchronize()
            t2 = time.time()
            e = (vis.to(torch.complex128).reshape(-1) - V).abs()
            got = dirty[ix[probe], iy[probe]].double().cpu().numpy()
            e2 = numpy.abs(got - exact)
            row = {"mode": mode, "eps": eps, "W": plan["W"], "sigma": plan["sigma"], "grid": plan["gx"],
                   "nplanes": plan["nplanes"], "invert_s": round(t1 - t0, 3), "predict_s": round(t2 - t1, 3),
                   "vis_max_rel": float(e.max() / V.abs().max()),
                   "vis_rms_rel": float(torch.sqrt((e ** 2).mean() / (V.abs() ** 2).mean())),
                   "dirty_max_rel": float(e2.max() / numpy.abs(exact).max()),
                   "dirty_rms_rel": float(numpy.sqrt((e2 ** 2).mean() / (exact ** 2).mean()))}
            print("[study]", row, file=sys.stderr, flush=True)
            rows.append(row)
            del dirty, vis, ms
    return {"study": "precision", "nvis": int(vtrue.size), "image": N, "cell": cell, "sources": nsrc,
            "rows": rows}


def main():
    args = parse()
    if args.sigma:
        os.environ["HIPNG_SIGMA"] = str(args.sigma)
    if args.study == "precision":
        print(json.dumps(precision_study()), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus %d is not supported "
                                   "(the sharding design is described in DESIGN.md)" % args.gpus}))
        sys.exit(2)
    import numpy
    import torch

    from rascil_amd import _lib, wgridder

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    wl = workload(args.config, args.nchan)
    prec = args.precision if args.config == "c1" else "fp64"
    nrow, nchan, npol, N = wl["uvw"].shape[0], len(wl["freq"]), wl["npol"], wl["npix"]
    nvis = nrow * nchan * npol
    print(f"[bench] {wl['name']}: {nrow} rows x {nchan} ch x {npol} pol = {nvis:.3e} vis, "
          f"{N}^2, precision {prec}", file=sys.stderr, flush=True)
    cdt = torch.complex128 if prec == "fp64" else torch.complex64
    rdt = torch.float64 if prec == "fp64" else torch.float32
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    uvw_d = torch.from_numpy(numpy.ascontiguousarray(wl["uvw"])).to(dev)
    vis_d = torch.randn((nrow, nchan, npol), dtype=cdt, device=dev, generator=g)
    model_d = torch.randn((npol, N, N), dtype=rdt, device=dev, generator=g)
    torch.cuda.synchronize()
    plan_i, plan_p = {}, {}

    def step(p, bc=None):
        # unit weights: wgt=None is the same sum without reading a weight array
        wgridder.ms2dirty(uvw_d, wl["freq"], vis_d, None, N, N, wl["cell"], wl["cell"], 0, 0,
                          wl["eps"], True, precision=p, plan_out=plan_i, batch_chunk=bc)
        wgridder.dirty2ms(uvw_d, wl["freq"], model_d, None, wl["cell"], wl["cell"], 0, 0,
                          wl["eps"], True, precision=p, plan_out=plan_p, batch_chunk=bc)

    for i in range(args.warmup):
        t = time.time()
        step(prec)
        torch.cuda.synchronize()
        print(f"[bench] warmup {i} {time.time() - t:.2f} s", file=sys.stderr, flush=True)
    _lib.profile_enable(True)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.time()
    e0.record()
    for i in range(args.steps):
        step(prec)
        print(f"[bench] step {i} done", file=sys.stderr, flush=True)
    e1.record()
    torch.cuda.synchronize()
    wall = time.time() - t0
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    ms_per_step = max(e0.elapsed_time(e1), wall * 1e3) / args.steps
    value = nvis / (ms_per_step * 1e-3) / 1e6

    csz = 16 if prec == "fp64" else 8
    kern = []
    for name, (tot, cnt) in prof.items():
        plan = plan_p if name in ("degrid", "screen_fill", "fft_c2c_forward") else plan_i
        b = algorithmic_bytes(name, plan, nrow * nchan, npol, csz)
        kern.append((tot, name, cnt, b))
    kern.sort(reverse=True)
    roof = None
    for tot, name, cnt, b in kern:
        if b is not None:
            avg_s = tot / cnt * 1e-3
            ach = b / avg_s / 1e9
            roof = {"kernel": name, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_launch_ms": round(tot / cnt, 4), "launches": cnt,
                    "algorithmic_bytes_per_launch": int(b)}
            break
    breakdown = {name: {"total_ms": round(tot, 2), "launches": cnt} for tot, name, cnt, _ in kern}

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mvis/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32", "fp32": "f32"}[prec], "data": "synthetic",
        "config": {"workload": wl["desc"], "nvis": nvis, "precision": prec,
                   "W": plan_i.get("W"), "sigma": plan_i.get("sigma"), "grid": plan_i.get("gx"),
                   "nplanes": plan_i.get("nplanes"), "tile": plan_i.get("tile")},
        "roofline": roof,
        "kernel_ms_total": breakdown,
    }

    if not args.no_accuracy:
        try:
            out["accuracy"] = accuracy_probe(wl, prec)
            print(f"[bench] accuracy {out['accuracy']}", file=sys.stderr, flush=True)
        except Exception as e:
            out["accuracy"] = {"error": repr(e)}

    if args.fp64_steps > 0 and prec != "fp64":
        del vis_d, model_d
        torch.cuda.empty_cache()
        _lib.load().hipng_release()
        vis64 = torch.randn((nrow, nchan, npol), dtype=torch.complex128, device=dev, generator=g)
        mod64 = torch.randn((npol, N, N), dtype=torch.float64, device=dev, generator=g)
        vis_d, model_d = vis64, mod64
        bc = 2 if nvis > 1e9 else None  # two polarisations per pass keeps FP64 buffers in HBM
        step("fp64", bc)
        torch.cuda.synchronize()
        t = time.time()
        for _ in range(args.fp64_steps):
            step("fp64", bc)
        torch.cuda.synchronize()
        ms64 = (time.time() - t) * 1e3 / args.fp64_steps
        out["fp64_plain"] = {"ms_per_step": round(ms64, 1), "value": round(nvis / (ms64 * 1e-3) / 1e6, 2),
                             "speedup_of_fp32_mode": round(ms64 / ms_per_step, 2),
                             "W": plan_i.get("W"), "grid": plan_i.get("gx"), "nplanes": plan_i.get("nplanes")}

    if not args.no_cpu:
        try:
            from oracle import wstack_cpu

            out["cpu_baseline"] = wstack_cpu.cpu_baseline(wl, plan_i, plan_p, args.cpu_seconds,
                                                          plans=cpu_plans(plan_i, npol, nrow * nchan))
        except Exception as e:  # the baseline leg must not hide the GPU result
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
