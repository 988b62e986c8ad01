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
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


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
                    help="pin the grid oversampling factor through the plan options (cost-model "
                         "calibration, not the bench line)")
    ap.add_argument("--tile", type=int, default=None, choices=[16, 32],
                    help="pin the uv tile (study option; the bench line uses the automatic plan)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a plain 1 GiB device copy 4 times (counter calibration for rocprof --pmc)")
    ap.add_argument("--study", default=None, choices=["precision", "autograd", "batched", "weighting", "gains",
                                                        "vis_passes", "dft"],
                    help="run a study instead of the bench line (prints one JSON line)")
    ap.add_argument("--rascil-steps", type=int, default=3,
                    help="also time invert_ng + predict_ng on device-resident columns (0: skip)")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="also time the FP64 plain mode for this many steps (0: skip)")
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


def algorithmic_work(name, plan, nvis, nitems, nb, csz, pc, dense_degrid=False):
    """Minimum work of one timed step for kernel class `name`: (FLOP, HBM bytes).

    FLOP follow the per-visibility counts of the gridding work model: 4 W^2 W_w (one complex x
    real FMA pair per cell and w plane) + 3 W c_ker (kernel taps, c_ker = 10) per visibility and
    direction; FFT 5 G^2 log2(G^2) per plane and polarisation. Bytes: each record is read once
    per pass of pc planes its w window meets, plus the occupied grid cells written (gridder:
    (plane, tile) pairs some footprint touches, from the plan) or the planes read (degridder,
    with the sorted accumulators read and written). The gridding of an invert is
    split between the sub-tile gatherer ("grid": planes written, records outside dense runs)
    and the dense-run gridder ("grid_dense": the plan's dense_items records); with
    `dense_degrid` the degridding of a predict is split the same way ("degrid": planes read,
    records outside dense runs; "degrid_dense": the dense_items records)."""
    G2 = plan["gx"] * plan["gy"]
    N2 = plan["nx"] * plan["ny"]
    P = plan["nplanes"]
    W = plan["W"]
    ww = W if plan["wstacking"] else 1
    rsz = 4 if csz == 8 else 8
    passes = (ww + pc - 1) / pc if plan["wstacking"] else 1  # passes an item takes part in
    dense = int(plan.get("dense_items", 0))
    per_vis = 4.0 * W * W * ww + 30.0 * W
    if name == "grid":
        # records and values read once per pass; the grid cells footprints touch written once
        # (at the tiles' granularity: occupied (plane, tile) pairs x tile^2, both gridding paths)
        items = nitems - dense
        occ = int(plan.get("occupied_tiles", 0)) * plan["tile"] ** 2
        return (items * nb * per_vis,
                items * passes * (4 + 3 * rsz + nb * csz) + occ * nb * csz)
    if name == "grid_dense":
        return (dense * nb * per_vis, dense * passes * (3 * rsz + nb * csz))
    if name == "degrid":
        items = nitems - dense if dense_degrid else nitems
        return (items * nb * per_vis, P * nb * G2 * csz + items * ww * (4 + 3 * rsz + 2 * nb * csz))
    if name == "degrid_dense":
        return (dense * nb * per_vis, dense * ww * (3 * rsz + 2 * nb * csz))
    if name.startswith("fft"):
        # rows transformed: the band of grid rows footprints reach (as the library sizes it)
        # and the image's ny columns on the other side, each read and written once
        U = plan["umax"] * plan["pixsize_x"] * plan["gx"]
        lo = math.floor(0.5 * plan["gx"] - U - 0.5 * W - 2.0)
        hi = math.ceil(0.5 * plan["gx"] + U + 0.5 * W + 2.0) + 1.0
        band = plan["gx"] if lo < 0 or hi > plan["gx"] else int(hi) - int(lo)
        rows = band * plan["gy"] + plan["ny"] * plan["gx"]
        return (nb * P * 5.0 * rows * math.log2(plan["gx"]), 2 * nb * P * rows * csz)
    if name == "screen_acc":
        return (None, P * nb * N2 * (csz + 16))
    if name == "screen_fill":
        return (None, P * nb * (N2 * 8 + G2 * csz))
    return (None, None)


def measured_traffic(kernel, plan, prec, nb):
    """HBM bytes per launch of kernel class `kernel` from the committed rocprofv3 --pmc
    summaries (profiles/*_traffic.json): FETCH_SIZE and WRITE_SIZE per dispatch, each scaled
    by what the counter reads on a plain device copy of known size. Only a summary measured
    on the same plan (precision, batch, kernel support, grid, planes, dense-run threshold) is
    used."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.json"))):
        try:
            with open(path) as f:
                rec = json.load(f)
        except (OSError, ValueError):
            continue
        pl = rec.get("plan", {})
        if (rec.get("kernel_class") == kernel and pl.get("precision") == prec and pl.get("nb") == nb
                and pl.get("W") == plan.get("W") and pl.get("grid") == plan.get("gx")
                and pl.get("nplanes") == plan.get("nplanes")
                and pl.get("dense_min", 0) == plan.get("dense_min", 0)):
            return rec["traffic_bytes_per_launch"], os.path.basename(path)
    return None, None


def cpu_plans(plan_i, npol, nitems):
    """Other oversampling choices for the CPU leg (it reports the fastest plan it tried)."""
    from rascil_amd import wgridder

    out = []
    for sg in (1.25, 2.0):
        if abs(sg - plan_i["sigma"]) < 1e-9:
            continue
        p = wgridder.plan_params(plan_i["umax"], plan_i["vmax"], plan_i["wmin"], plan_i["wmax"],
                                 plan_i["nx"], plan_i["ny"], plan_i["pixsize_x"], plan_i["pixsize_y"],
                                 plan_i["epsilon"], True, "fp64", npol, nitems,
                                 plan_opts={"sigma": sg}).as_dict()
        out.append((p, p))
    return out


def accuracy_at_timed_plan(wl, prec, uvw_d, ms_buf, wgt_d, timed_plans, nsrc=20, seed=7, chunk_rows=1 << 20):
    """Max / RMS relative error of predict and invert against FP64 direct sums over the WHOLE
    workload, through the same one-shot calls as the timed step (same uvw, batch, weights,
    epsilon and precision, hence the same plan, which is compared with the timed one):

    * predict of `nsrc` pixel-centred sources, polarisation b scaled by (1, 0.5, -0.25,
      0.125)[b], against the FP64 HIP DFT of flux / n (the w-stacking convention) at every
      visibility;
    * invert of those exact visibilities (rounded to the input type, written into `ms_buf`)
      with the timed step's weights, at the source pixels, against the FP64 adjoint sum
      Re sum_i w_i V_i exp(+2 pi i phi_i) / n over every visibility (torch, FP64, on the GPU).
    """
    import numpy
    import torch

    from rascil_amd import wgridder
    from rascil_amd.dft import dft_device

    t_start = time.time()
    dev = uvw_d.device
    rng = numpy.random.default_rng(seed)
    freq, N, cell, eps = numpy.asarray(wl["freq"], "f8"), wl["npix"], wl["cell"], wl["eps"]
    nrow, nchan, nb = ms_buf.shape
    rdt = torch.float64 if prec == "fp64" else torch.float32
    ix = rng.integers(N // 8, N - N // 8, nsrc)
    iy = rng.integers(N // 8, N - N // 8, nsrc)
    flux = rng.uniform(0.1, 1.0, nsrc)
    scale = numpy.array([1.0, 0.5, -0.25, 0.125])[:nb]
    model = torch.zeros((nb, N, N), dtype=rdt, device=dev)
    ixt, iyt = torch.tensor(ix, device=dev), torch.tensor(iy, device=dev)
    for b in range(nb):
        model[b].index_put_((ixt, iyt), torch.tensor(flux * scale[b], dtype=rdt, device=dev), accumulate=True)
    x = (ix - N // 2) * cell
    y = (iy - N // 2) * cell
    n = numpy.sqrt(1 - x * x - y * y)
    plan_p = {}
    vis = wgridder.dirty2ms(uvw_d, freq, model, None, cell, cell, 0, 0, eps, True, precision=prec, plan_out=plan_p)
    dc = torch.tensor(numpy.stack([-x, y, n - 1], axis=1), dtype=torch.float64, device=dev)
    fl = torch.tensor(flux / n, dtype=torch.complex128, device=dev).reshape(nsrc, 1, 1).expand(nsrc, nchan, 1).contiguous()
    sgn = torch.tensor([-1.0, 1.0, -1.0], dtype=torch.float64, device=dev)
    fc = torch.tensor(freq / 299792458.0, dtype=torch.float64, device=dev)
    sc = torch.tensor(scale, dtype=torch.complex128, device=dev)
    pe_max = torch.zeros(nb, dtype=torch.float64, device=dev)
    pe_sq = torch.zeros(nb, dtype=torch.float64, device=dev)
    pr_max = torch.zeros(nb, dtype=torch.float64, device=dev)
    pr_sq = torch.zeros(nb, dtype=torch.float64, device=dev)
    for r0 in range(0, nrow, chunk_rows):
        r1 = min(nrow, r0 + chunk_rows)
        uvl = ((uvw_d[r0:r1] * sgn)[:, None, :] * fc[None, :, None]).contiguous()
        ref = dft_device(dc, fl, uvl.reshape(1, r1 - r0, nchan, 3))[0, :, :, 0]
        refb = ref[..., None] * sc  # [rows, nchan, nb]
        err = (vis[r0:r1].to(torch.complex128) - refb).abs()
        pe_max = torch.maximum(pe_max, err.amax(dim=(0, 1)))
        pe_sq += (err ** 2).sum(dim=(0, 1))
        ra = refb.abs()
        pr_max = torch.maximum(pr_max, ra.amax(dim=(0, 1)))
        pr_sq += (ra ** 2).sum(dim=(0, 1))
        ms_buf[r0:r1] = refb.to(ms_buf.dtype)
        del uvl, ref, refb, err, ra
    del vis
    torch.cuda.empty_cache()
    plan_i = {}
    dirty = wgridder.ms2dirty(uvw_d, freq, ms_buf, wgt_d, N, N, cell, cell, 0, 0, eps, True, precision=prec,
                              plan_out=plan_i)
    got = dirty[:, ixt, iyt].double()  # [nb, nsrc]
    del dirty
    exact = torch.zeros((nb, nsrc), dtype=torch.float64, device=dev)
    dcs = dc.cpu().numpy()
    for r0 in range(0, nrow, chunk_rows):
        r1 = min(nrow, r0 + chunk_rows)
        uvl = (uvw_d[r0:r1] * sgn)[:, None, :] * fc[None, :, None]
        V = ms_buf[r0:r1].to(torch.complex128)
        if wgt_d is not None:
            V = V * wgt_d[r0:r1].double()
        for k in range(nsrc):
            ph = uvl[..., 0] * dcs[k, 0] + uvl[..., 1] * dcs[k, 1] + uvl[..., 2] * dcs[k, 2]
            e = torch.polar(torch.ones_like(ph), 2 * numpy.pi * ph)
            exact[:, k] += (V * e[..., None]).real.sum(dim=(0, 1))
        del uvl, V
    exact = exact / torch.tensor(n, dtype=torch.float64, device=dev)
    e2 = (got - exact).abs()
    nv = nrow * nchan

    def _res(emax, esq, rmax, rsq):
        return {"max_rel": float(emax.max() / rmax.max()), "rms_rel": float(torch.sqrt(esq.sum() / rsq.sum())),
                "per_pol_max_rel": [float(a) for a in (emax / rmax).cpu()]}

    keys = ("W", "sigma", "gx", "gy", "nplanes", "tile", "dense_min", "dense_items")
    same = all(plan_i.get(k) == timed_plans[0].get(k) for k in keys) and all(
        plan_p.get(k) == timed_plans[1].get(k) for k in keys)
    out = {"visibilities": int(nv * nb), "rows": int(nrow), "sources": nsrc, "pol_scales": scale.tolist(),
           "plan": {k: plan_i.get(k) for k in keys}, "plan_is_timed_plan": bool(same),
           "predict": _res(pe_max, pe_sq, pr_max, pr_sq),
           "invert": {"max_rel": float(e2.max() / exact.abs().max()),
                      "rms_rel": float(torch.sqrt((e2 ** 2).mean() / (exact ** 2).mean())),
                      "per_pol_max_rel": [float(a) for a in (e2.amax(dim=1) / exact.abs().amax(dim=1)).cpu()]},
           "seconds": round(time.time() - t_start, 1)}
    out["meets_epsilon"] = bool(out["predict"]["max_rel"] <= eps and out["invert"]["max_rel"] <= eps and same)
    return out


def rascil_level_leg(wl, prec, uvw_d, steps, wgridder_ms):
    """invert_ng + predict_ng (the RASCIL surface) on a DeviceBlockVisibility of the workload:
    complex64 visibilities and float32 imaging weights resident in GPU memory, a linear
    polarisation frame imaged as Stokes IQUV (the fused gather / unsort do the conversion and
    the weighting), the image on the device. The first call builds and keeps the geometry
    (sort, records, plans); the timed steps then reuse it, as a major cycle does."""
    import numpy
    import torch

    from rascil_amd import DeviceBlockVisibility, PolarisationFrame, create_image, invert_ng, ng, predict_ng
    from rascil_amd.data_models import Direction

    dev = uvw_d.device
    nrow, nchan, N, cell = uvw_d.shape[0], len(wl["freq"]), wl["npix"], wl["cell"]
    g = torch.Generator(device=dev)
    g.manual_seed(99)
    vis = torch.randn((1, nrow, nchan, 4), dtype=torch.complex64, device=dev, generator=g)
    wt = torch.empty((1, nrow, nchan, 4), dtype=torch.float32, device=dev).uniform_(0.5, 1.5, generator=g)
    pc = Direction(0.0, numpy.radians(-30.0))
    dbv = DeviceBlockVisibility(vis, uvw_d.reshape(1, nrow, 3), wl["freq"], PolarisationFrame("linear"), pc,
                                weight=wt)
    model = create_image(N, cell, pc, [float(numpy.mean(wl["freq"]))], PolarisationFrame("stokesIQUV"))
    model.pixels = torch.randn((1, 4, N, N), dtype=torch.float32, device=dev, generator=g)
    kw = dict(epsilon=wl["eps"], precision=prec)
    ng.clear_geometry_cache()
    torch.cuda.synchronize()
    t = time.time()
    invert_ng(dbv, model, **kw)
    pred = predict_ng(dbv, model, **kw)
    torch.cuda.synchronize()
    first = (time.time() - t) * 1e3
    del pred
    torch.cuda.synchronize()
    t = time.time()
    for i in range(steps):
        im, sumwt = invert_ng(dbv, model, **kw)
        pred = predict_ng(dbv, model, **kw)
        del im, pred
        print(f"[bench] rascil-level step {i} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    ms = (time.time() - t) * 1e3 / steps
    ng.clear_geometry_cache()
    nvis = nrow * nchan * 4
    return {"api": "invert_ng + predict_ng on a DeviceBlockVisibility (complex64 / float32 columns, linear -> "
                   "stokesIQUV, image on the device, geometry kept between calls)",
            "ms_per_step": round(ms, 1), "value": round(nvis / (ms * 1e-3) / 1e6, 2), "steps": steps,
            "first_call_ms": round(first, 1), "ratio_to_wgridder_level": round(ms / wgridder_ms, 3)}


def path_roofline(plan_i, plan_p, nrow, nchan, nb, prec, with_wgt, ms_per_step):
    """Whole-path floor of one invert + predict by the work model of the scope table:
    T_roof = max(HBM bytes / 8 TB/s, FLOP / peak) with
    FLOP = per direction N_vis (4 W^2 W_w + 3 W 10) + N_pol N_planes 5 G^2 log2 G^2, and bytes =
    visibilities read (+ weights) and written, uvw read per direction, the sorted index
    written and read per direction (one per row-channel item, shared by the polarisations),
    N_pol N_planes G^2 b (2 p_fft + 2) per direction for the plane FFTs (p_fft = 2) and the
    screen, plus the image written / read."""
    bv = 16 if prec == "fp64" else 8
    bw = 8 if prec == "fp64" else 4
    br = 8 if prec == "fp64" else 4
    items = nrow * nchan
    nvis = items * nb
    flop = 0.0
    nbytes = nvis * (bv + (bw if with_wgt else 0)) + nvis * bv + 2 * nrow * 24 + 2 * 2 * items * 4
    for P in (plan_i, plan_p):
        W = P["W"]
        ww = W if P["wstacking"] else 1
        G2 = P["gx"] * P["gy"]
        flop += nvis * (4.0 * W * W * ww + 30.0 * W) + nb * P["nplanes"] * 5.0 * G2 * math.log2(G2)
        nbytes += nb * P["nplanes"] * G2 * bv * (2 * 2 + 2) + nb * P["nx"] * P["ny"] * br
    peak = (FP64_PEAK_TFLOPS if prec == "fp64" else FP32_PEAK_TFLOPS) * 1e12
    t_hbm = nbytes / (HBM_PEAK_GBS * 1e9)
    t_flop = flop / peak
    t_roof = max(t_hbm, t_flop)
    return {"bound": "hbm" if t_hbm >= t_flop else "flop", "hbm_bytes": int(nbytes), "flop": int(flop),
            "t_hbm_ms": round(t_hbm * 1e3, 1), "t_flop_ms": round(t_flop * 1e3, 1),
            "t_roof_ms": round(t_roof * 1e3, 1), "ms_per_step": round(ms_per_step, 1),
            "frac": round(t_roof * 1e3 / ms_per_step, 4)}


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


def autograd_study(seed=11, nsrc=500, descent_steps=50):
    """Differentiable predict at 2048^2 (cell 2e-5 rad): ~1M rows x 4 channels from 256
    stations, 256 per-antenna complex gains (|g| = 1 +- 0.1, phase +-30 deg), V_obs = FP64
    DFT of power-law sources + noise 0.01. Loss = sum |V_obs - g_i conj(g_j) predict(I)|^2.
    Image and gain gradients are checked by central finite differences (exact for the
    image: the loss is quadratic in it) and the image gradient against the FP64 adjoint
    sum; then gradient descent on the gains."""
    import numpy
    import torch

    from rascil_amd import simulation, wgridder
    from rascil_amd.autograd import apply_gains, predict
    from rascil_amd.dft import dft_kernel

    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(seed)
    nant = 256
    ants = simulation.core_plus_arms_stations(nant, 0.7, 500.0, 1000.0, 40000.0, rng)
    a1, a2 = simulation.baseline_indices(nant, autocorrelations=False)
    ntimes = int(numpy.ceil(1_000_000 / len(a1)))
    dec = numpy.radians(-30.0)
    uvw = simulation.baseline_uvw(ants, numpy.linspace(-2, 2, ntimes) * numpy.pi / 12, dec,
                                  autocorrelations=False).reshape(-1, 3)
    ant1, ant2 = numpy.tile(a1, ntimes), numpy.tile(a2, ntimes)
    freq = numpy.array([100e6, 110e6, 120e6, 130e6])
    N, cell = 2048, 2e-5
    # sky at pixel centres; V_obs = DFT of flux / n (the gridder's w-stacking convention)
    ix = rng.integers(N // 8, N - N // 8, nsrc)
    iy = rng.integers(N // 8, N - N // 8, nsrc)
    flux = simulation.powerlaw_fluxes(nsrc, -1.6, 1e-3, 1.0, rng)
    x, y = (ix - N // 2) * cell, (iy - N // 2) * cell
    n = numpy.sqrt(1 - x * x - y * y)
    uvl = ((uvw * numpy.array([-1.0, 1.0, -1.0]))[:, None, :] * (freq / 299792458.0)[None, :, None])[None]
    vtrue = dft_kernel(numpy.stack([-x, y, n - 1], 1), (flux / n)[:, None, None] * numpy.ones((1, 4, 1)), uvl)
    vtrue = vtrue[0, :, :, 0]
    gtrue = rng.uniform(0.9, 1.1, nant) * numpy.exp(1j * numpy.radians(rng.uniform(-30, 30, nant)))
    vobs = gtrue[ant1, None] * numpy.conj(gtrue[ant2, None]) * vtrue
    vobs = vobs + 0.01 * (rng.normal(size=vobs.shape) + 1j * rng.normal(size=vobs.shape))
    out = {"study": "autograd", "rows": int(uvw.shape[0]), "channels": 4, "image": N, "stations": nant,
           "sources": nsrc}
    uvw_d = torch.tensor(uvw, device=dev)
    i1 = torch.tensor(ant1, device=dev)
    i2 = torch.tensor(ant2, device=dev)
    model = numpy.zeros((N, N))
    numpy.add.at(model, (ix, iy), flux)
    for prec, eps in (("fp64", 1e-10), ("mixed", 1e-5)):
        rd = torch.float64 if prec == "fp64" else torch.float32
        cd = torch.complex64 if prec == "mixed" else torch.complex128
        t = time.time()
        geom = wgridder.Geometry(uvw_d, freq, N, N, cell, cell, eps, True, prec)
        t_geom = time.time() - t
        vo = torch.tensor(vobs, dtype=torch.complex128, device=dev)
        img0 = torch.tensor(model * 0.9, dtype=rd, device=dev)
        g0 = torch.tensor(gtrue * (1 + 0.05 * rng.normal(size=nant)), dtype=cd, device=dev)

        def loss(img, g):
            v = predict(img, geometry=geom).to(torch.complex128)
            r = vo - apply_gains(v, g.to(torch.complex128), i1, i2)
            return (r.abs() ** 2).sum()

        img = img0.clone().requires_grad_(True)
        g = g0.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t = time.time()
        L = loss(img, g)
        L.backward()
        torch.cuda.synchronize()
        t_fb = time.time() - t
        gi, gg = img.grad.detach(), g.grad.detach()
        errs_i, errs_g = [], []
        h, hg = 1.0, (1e-5 if prec == "fp64" else 1e-2)
        with torch.no_grad():
            for k in range(4):
                px, py = int(ix[k]), int(iy[k])
                ip, im_ = img0.clone(), img0.clone()
                ip[px, py] += h
                im_[px, py] -= h
                fd = (loss(ip, g0) - loss(im_, g0)).item() / (2 * h)
                errs_i.append(abs(fd - gi[px, py].item()) / max(1.0, abs(fd)))
            for k in (0, 7):
                for dirn in (1.0, 1j):
                    gp, gm = g0.clone(), g0.clone()
                    gp[k] += hg * dirn
                    gm[k] -= hg * dirn
                    fd = (loss(img0, gp) - loss(img0, gm)).item() / (2 * hg)
                    an = gg[k].real.item() if dirn == 1.0 else gg[k].imag.item()
                    errs_g.append(abs(fd - an) / max(1.0, abs(fd)))
            # image gradient vs the FP64 adjoint sum at the source pixels
            v = predict(img0, geometry=geom).to(torch.complex128)
            gprod = g0.to(torch.complex128)[i1] * torch.conj(g0.to(torch.complex128))[i2]
            back = (-2.0 * torch.conj(gprod)[:, None] * (vo - gprod[:, None] * v)).reshape(-1)
            U = torch.tensor(uvl[0].reshape(-1, 3), device=dev)
            ad = []
            for k in range(8):
                ph = U[:, 0] * (-x[k]) + U[:, 1] * y[k] + U[:, 2] * (n[k] - 1)
                ad.append(float((back * torch.exp(2j * numpy.pi * ph.to(torch.complex128))).real.sum()) / n[k])
            ad = numpy.array(ad)
            got = numpy.array([gi[int(ix[k]), int(iy[k])].item() for k in range(8)])
            adj_err = float(numpy.max(numpy.abs(got - ad)) / numpy.max(numpy.abs(ad)))
        # gradient descent on the gains (image fixed at the true model)
        imgt = torch.tensor(model, dtype=rd, device=dev)
        gd = torch.ones(nant, dtype=cd, device=dev).requires_grad_(True)
        opt = torch.optim.Adam([gd], lr=0.02)
        hist = []
        torch.cuda.synchronize()
        t = time.time()
        for it in range(descent_steps):
            opt.zero_grad()
            Lg = loss(imgt, gd)
            Lg.backward()
            opt.step()
            hist.append(Lg.item())
        torch.cuda.synchronize()
        t_desc = time.time() - t
        out[prec] = {"epsilon": eps, "plan": {k: geom.plan[k] for k in ("W", "sigma", "gx", "nplanes")},
                     "geometry_s": round(t_geom, 3), "forward_backward_s": round(t_fb, 3),
                     "fd_image_max_rel": max(errs_i), "fd_gain_max_rel": max(errs_g),
                     "image_grad_vs_fp64_adjoint_max_rel": adj_err,
                     "descent": {"steps": descent_steps, "loss_first": hist[0], "loss_last": hist[-1],
                                 "seconds_per_step": round(t_desc / descent_steps, 4)}}
        print("[study]", prec, out[prec], file=sys.stderr, flush=True)
        geom.close()
    return out


def batched_study(njobs=300, seed=5):
    """Many independent snapshot jobs of uneven size: 300 jobs, rows log-uniform 1k-2M,
    images from {256..4096}^2 with the cell matched to each job's longest baseline,
    1-16 channels, 2-20 min windows, dec in [-80, +10] deg,
    epsilon 1e-4 (mixed precision). Serial loop of one-shot invert+predict calls vs
    prebuilt geometries whose invert+predict are captured in hipGraphs and replayed."""
    import numpy
    import torch

    from rascil_amd import simulation, wgridder

    dev = torch.device("cuda", 0)
    rng = numpy.random.default_rng(seed)
    ants = simulation.core_plus_arms_stations(512, 0.7, 500.0, 1000.0, 40000.0, rng)
    a1, a2 = simulation.baseline_indices(512, autocorrelations=False)
    nbl = len(a1)
    jobs = []
    total = 0
    for jb in range(njobs):
        rows = int(numpy.exp(rng.uniform(numpy.log(1e3), numpy.log(2e6))))
        nch = int(rng.integers(1, 17))
        N = int(rng.choice([256, 512, 1024, 2048, 4096]))
        dec = numpy.radians(rng.uniform(-80, 10))
        win = rng.uniform(2, 20) / 60.0 * numpy.pi / 12
        nt = max(1, rows // nbl + 1)
        ha = numpy.linspace(-win / 2, win / 2, nt)
        sel = rng.choice(nbl, min(rows, nbl), replace=False)
        uvw = simulation.baseline_uvw(ants, ha, dec, autocorrelations=False)[:, sel].reshape(-1, 3)[:rows]
        freq = numpy.linspace(100e6, 164e6, nch)
        umax = numpy.abs(uvw[:, :2]).max() * freq.max() / 299792458.0
        cell = 1.0 / (2.2 * umax)  # 2.2 pixels across the finest fringe
        jobs.append(dict(uvw=uvw, freq=freq, N=N, cell=cell, nvis=rows * nch))
        total += rows * nch
    eps = 1e-4
    out = {"study": "batched", "jobs": len(jobs), "nvis": int(total)}
    # serial loop of one-shot calls
    lat = []
    torch.cuda.synchronize()
    t0 = time.time()
    for j in jobs:
        u = torch.tensor(j["uvw"], device=dev)
        ms = torch.ones((len(j["uvw"]), len(j["freq"])), dtype=torch.complex64, device=dev)
        img = torch.zeros((j["N"], j["N"]), dtype=torch.float32, device=dev)
        t = time.time()
        wgridder.ms2dirty(u, j["freq"], ms, None, j["N"], j["N"], j["cell"], j["cell"], 0, 0, eps, True,
                          precision="mixed")
        wgridder.dirty2ms(u, j["freq"], img, None, j["cell"], j["cell"], 0, 0, eps, True, precision="mixed")
        torch.cuda.synchronize()
        lat.append(time.time() - t)
    t_serial = time.time() - t0
    out["serial"] = {"seconds": round(t_serial, 3), "mvis_per_s": round(total / t_serial / 1e6, 2),
                     "p50_ms": round(1e3 * float(numpy.percentile(lat, 50)), 2),
                     "p99_ms": round(1e3 * float(numpy.percentile(lat, 99)), 2)}
    print("[study] serial", out["serial"], file=sys.stderr, flush=True)
    # geometries built once (in groups), invert+predict captured in one hipGraph per group
    lat = []
    t_build = 0.0
    t_run = 0.0
    group = 24
    for g0 in range(0, len(jobs), group):
        grp = jobs[g0:g0 + group]
        t = time.time()
        items = []
        for j in grp:
            geom = wgridder.Geometry(j["uvw"], j["freq"], j["N"], j["N"], j["cell"], j["cell"], eps, True, "mixed")
            ms = torch.ones((len(j["uvw"]), len(j["freq"]), 1), dtype=torch.complex64, device=dev)
            img = torch.zeros((1, j["N"], j["N"]), dtype=torch.float32, device=dev)
            d = torch.empty((1, j["N"], j["N"]), dtype=torch.float32, device=dev)
            v = torch.empty_like(ms)
            items.append((geom, ms, img, d, v))
        torch.cuda.synchronize()
        t_build += time.time() - t
        st = torch.cuda.Stream()
        st.wait_stream(torch.cuda.current_stream())
        graphs = []
        with torch.cuda.stream(st):
            for geom, ms, img, d, v in items:  # warm-up outside capture
                geom.ms2dirty(ms, out=d)
                geom.dirty2ms(img, out=v)
        torch.cuda.current_stream().wait_stream(st)
        torch.cuda.synchronize()
        for geom, ms, img, d, v in items:
            gph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(gph):
                geom.ms2dirty(ms, out=d)
                geom.dirty2ms(img, out=v)
            graphs.append(gph)
        torch.cuda.synchronize()
        t = time.time()
        for gph in graphs:
            t1 = time.time()
            gph.replay()
            torch.cuda.synchronize()
            lat.append(time.time() - t1)
        t_run += time.time() - t
        for geom, *_ in items:
            geom.close()
        del items, graphs
        torch.cuda.empty_cache()
    out["graph_replay"] = {"seconds": round(t_run, 3), "mvis_per_s": round(total / t_run / 1e6, 2),
                           "p50_ms": round(1e3 * float(numpy.percentile(lat, 50)), 2),
                           "p99_ms": round(1e3 * float(numpy.percentile(lat, 99)), 2),
                           "geometry_build_seconds": round(t_build, 3)}
    print("[study] graph", out["graph_replay"], file=sys.stderr, flush=True)
    # plan options tuned per job class, jobs' graphs replayed concurrently over 4 streams
    from rascil_amd.batch import AutoTuner, Batch, class_table

    tuner = AutoTuner()
    t = time.time()
    for j in jobs:
        tuner.tune(j["uvw"], j["freq"], j["N"], j["cell"], eps, "mixed")
    t_tune = time.time() - t
    lat, t_run, t_build = [], 0.0, 0.0
    for g0 in range(0, len(jobs), group):
        grp = [dict(uvw=j["uvw"], freq=j["freq"], npix=j["N"], cell=j["cell"], nvis=j["nvis"])
               for j in jobs[g0:g0 + group]]
        t = time.time()
        b = Batch(grp, eps, "mixed", tuner=tuner, nstreams=4)
        t_build += time.time() - t
        r = b.run()
        t_run += r["seconds"]
        lat += r["latency_ms"]
        b.close()
        del b
        torch.cuda.empty_cache()
    out["batch_tuned_4_streams"] = {"seconds": round(t_run, 3), "mvis_per_s": round(total / t_run / 1e6, 2),
                                    "p50_ms": round(float(numpy.percentile(lat, 50)), 2),
                                    "p99_ms": round(float(numpy.percentile(lat, 99)), 2),
                                    "geometry_build_seconds": round(t_build, 3), "tuning_seconds": round(t_tune, 1),
                                    "classes": len(tuner.table), "tuned": class_table(tuner)}
    print("[study] batch", {k: v for k, v in out["batch_tuned_4_streams"].items() if k != "tuned"},
          file=sys.stderr, flush=True)
    return out


def weighting_study(reps=5, seed=9, cpu_fracs=(0.0025, 0.005)):
    """Imaging weights (uniform and robust) at the production shape: 16,744,448 rows x 64
    channels x 4 polarisations of FP64 imaging weights (seeded uniform 0.5-1.5), density on the
    8192^2 uv grid of the 1.1e-5 rad image (one image channel). Inputs resident in HBM; each
    timed call starts from the same weights (restored outside the timed region). Throughput in
    Mvis/s (weights per second); HBM view of the whole call against its algorithmic bytes:
    weights read by the density pass and read + written by the reweighting pass (24 B per
    visibility), uvw read by both passes (48 B per row), the density grid cleared, summed and
    read (3 x 8 B per cell). CPU baseline: the oracle's numpy restatement on two row
    subsamples through the full grid, extrapolated linearly to all rows."""
    import ctypes

    import numpy
    import torch

    from rascil_amd import _lib, simulation

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    lib = _lib.load()
    ntimes = 128
    uvw, freq, _ = simulation.config1_geometry(ntimes=ntimes)
    nrow, nchan, npol, N, cell = uvw.shape[0], len(freq), 4, 8192, 1.1e-5
    nvis = nrow * nchan * npol
    cell_u, cell_v = -1.0 / (N * cell), 1.0 / (N * cell)  # RA axis: negative cdelt
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    uvw_d = torch.from_numpy(numpy.ascontiguousarray(uvw)).to(dev)
    wt0 = torch.rand((nrow, nchan, npol), dtype=torch.float64, device=dev, generator=g) + 0.5
    wt = torch.empty_like(wt0)
    grid = torch.empty((1, N, N, npol), dtype=torch.float64, device=dev)
    v2i = numpy.zeros(nchan, numpy.int32)
    fq = numpy.ascontiguousarray(freq, "f8")
    st = torch.cuda.current_stream(dev)
    out = {"study": "weighting", "nvis": int(nvis), "rows": int(nrow), "channels": nchan, "pols": npol,
           "grid": [1, N, N, npol], "cell_rad": cell, "data": "synthetic"}

    def call(mode, rob):
        _lib.check(lib.hipng_weight_visibility(
            ctypes.c_void_p(uvw_d.data_ptr()), _lib.dptr(fq), ntimes, nrow // ntimes, nchan, npol,
            v2i.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1, N, N, cell_u, cell_v, mode, rob,
            ctypes.c_void_p(wt.data_ptr()), ctypes.c_void_p(grid.data_ptr()), ctypes.c_void_p(st.cuda_stream)))

    algo = 24.0 * nvis + 48.0 * nrow + 3 * 8.0 * grid.numel()
    for name, mode, rob in (("uniform", 1, 0.0), ("robust", 2, 0.5)):
        wt.copy_(wt0)
        call(mode, rob)  # warm-up
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        ms = []
        for _ in range(reps):
            wt.copy_(wt0)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            call(mode, rob)
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        prof = _lib.profile_read()
        _lib.profile_enable(False)
        t = float(numpy.median(ms))
        res = {"ms_per_call": round(t, 3), "value": round(nvis / (t * 1e-3) / 1e6, 1), "unit": "Mvis/s",
               "kernel_ms": {k: round(tot / cnt, 3) for k, (tot, cnt) in prof.items()},
               "roofline": {"bound": "hbm", "achieved": round(algo / (t * 1e-3) / 1e9, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(algo / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "algorithmic_bytes_per_call": int(algo)},
               "sum_weights": float(wt.sum()), "zero_weights": int((wt == 0).sum())}
        out[name] = res
        print("[study] weighting", name, res, file=sys.stderr, flush=True)
    # CPU baseline: the oracle on row subsamples (time = fixed grid cost + per-row cost)
    from oracle import weighting as ow

    rng = numpy.random.default_rng(seed)
    pts = []
    for f in cpu_fracs:
        rows = numpy.sort(rng.choice(nrow, int(nrow * f), replace=False))
        w = rng.uniform(0.5, 1.5, (len(rows), nchan, npol))
        t0 = time.time()
        ow.weight(uvw[rows], freq, w, v2i.astype(int), 1, N, N, numpy.degrees(-cell), numpy.degrees(cell),
                  "uniform")
        pts.append((len(rows), time.time() - t0))
    (r1, t1), (r2, t2) = pts
    per_row = (t2 - t1) / (r2 - r1)
    t_full = t1 + per_row * (nrow - r1)
    out["cpu_baseline"] = {"value": round(nvis / t_full / 1e6, 3), "unit": "Mvis/s", "cores": 1, "kind": "port",
                           "sample": (f"numpy oracle (uniform) on {r1} and {r2} of {nrow} rows through the full "
                                      f"8192^2 grid ({t1:.1f} s, {t2:.1f} s), extrapolated linearly to all rows")}
    out["speedup_vs_cpu_uniform"] = round(out["uniform"]["value"] / out["cpu_baseline"]["value"], 1)
    return out


def gains_study(reps=5, seed=13, cpu_times=(1, 2)):
    """Gain application (apply_gaintable) at the production shape: 128 times x 130,816
    baselines x 64 channels of linear-polarisation complex128 visibilities (4.29e9, 68.6 GB)
    with f64 weights, a per-channel (B) gain table of 2x2 Jones matrices per time, antenna
    (512) and channel; forward and inverse. Inputs resident in HBM; the host's time matching
    is outside the timed call. HBM view against the algorithmic bytes of a call: every
    visibility read and written (128 B), weights untouched (2x2 gains, no flags), the gain
    table read and its applied form written and read. CPU baseline: the numpy oracle on the
    visibilities of one and two times, extrapolated linearly to all times."""
    import ctypes

    import numpy
    import torch

    from rascil_amd import _lib, simulation

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    lib = _lib.load()
    ntimes, nant, nchan, npol = 128, 512, 64, 4
    a1, a2 = simulation.baseline_indices(nant, autocorrelations=False)
    nbl = len(a1)
    nvis = ntimes * nbl * nchan * npol
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    vis = torch.randn((ntimes, nbl, nchan, npol), dtype=torch.complex128, device=dev, generator=g)
    wt = torch.ones((ntimes, nbl, nchan, npol), dtype=torch.float64, device=dev)
    rng = numpy.random.default_rng(seed)
    gshape = (ntimes, nant, nchan, 2, 2)
    gain = numpy.zeros(gshape, complex)
    for k in range(2):
        gain[..., k, k] = numpy.exp(1j * 0.1 * rng.normal(size=gshape[:3]))
    gain[..., 0, 1] = 0.01 * rng.normal(size=gshape[:3])
    gain[..., 1, 0] = 0.01 * rng.normal(size=gshape[:3])
    gw = numpy.ones(gshape)
    dg = torch.from_numpy(gain).to(dev)
    dgw = torch.from_numpy(gw).to(dev)
    d1 = torch.from_numpy(a1.astype(numpy.int32)).to(dev)
    d2 = torch.from_numpy(a2.astype(numpy.int32)).to(dev)
    t2g = torch.arange(ntimes, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream(dev)
    out = {"study": "gains", "nvis": int(nvis), "times": ntimes, "baselines": nbl, "channels": nchan,
           "pols": npol, "jones": "B (2x2 per time, antenna, channel)", "data": "synthetic"}

    def call(inverse):
        _lib.check(lib.hipng_apply_gaintable(
            ctypes.c_void_p(vis.data_ptr()), ctypes.c_void_p(wt.data_ptr()), None, None, ntimes, nbl, nchan, npol,
            ctypes.c_void_p(d1.data_ptr()), ctypes.c_void_p(d2.data_ptr()), ctypes.c_void_p(dg.data_ptr()),
            ctypes.c_void_p(dgw.data_ptr()), ntimes, nant, nchan, 2, ctypes.c_void_p(t2g.data_ptr()),
            1 if inverse else 0, ctypes.c_void_p(st.cuda_stream)))

    algo = 128.0 * (nvis // npol) + 3 * gain.nbytes + 4 * ntimes * nant * nchan
    for name, inverse in (("forward", False), ("inverse", True)):
        call(inverse)
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        ms = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            call(inverse)
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        prof = _lib.profile_read()
        _lib.profile_enable(False)
        t = float(numpy.median(ms))
        res = {"ms_per_call": round(t, 3), "value": round(nvis / (t * 1e-3) / 1e6, 1), "unit": "Mvis/s",
               "kernel_ms": {k: round(tot / cnt, 3) for k, (tot, cnt) in prof.items()},
               "roofline": {"bound": "hbm", "achieved": round(algo / (t * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(algo / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "algorithmic_bytes_per_call": int(algo)}}
        out[name] = res
        print("[study] gains", name, res, file=sys.stderr, flush=True)
    # CPU baseline: the oracle on the visibilities of the first one and two times
    from oracle import gains as og

    pts = []
    vh = vis[:max(cpu_times)].cpu().numpy()
    for k in cpu_times:
        t0 = time.time()
        og.apply_gaintable(vh[:k], numpy.ones(vh[:k].shape), numpy.zeros(vh[:k].shape), numpy.arange(k, dtype=float),
                           a1, a2, gain[:k], gw[:k], numpy.arange(k, dtype=float), numpy.ones(k))
        pts.append((k, time.time() - t0))
    (k1, s1), (k2, s2) = pts
    t_full = s1 + (s2 - s1) / (k2 - k1) * (ntimes - k1)
    out["cpu_baseline"] = {"value": round(nvis / t_full / 1e6, 3), "unit": "Mvis/s", "cores": 1, "kind": "port",
                           "sample": (f"numpy oracle (forward) on the visibilities of {k1} and {k2} of {ntimes} "
                                      f"times ({s1:.1f} s, {s2:.1f} s), extrapolated linearly to all times")}
    out["speedup_vs_cpu_forward"] = round(out["forward"]["value"] / out["cpu_baseline"]["value"], 1)
    return out


def vis_passes_study(reps=5, seed=17):
    """The fused visibility passes of invert_ng / predict_ng at the production shape:
    16,744,448 rows x 64 channels of linear-polarisation complex128 visibilities (4.29e9) with
    int32 flags and f64 imaging weights; the image centre 0.5 deg from the phase centre, image
    frame stokesIQUV; gridder precision mixed (complex64 / f32). Prologue: flag mask, phase
    shift, linear -> IQUV, cast and layout for the gridder; epilogue: IQUV -> linear and the
    inverse phase shift of complex64 degridder output. Inputs resident in HBM. Algorithmic
    bytes per 4-pol item: prologue 64 (vis) + 16 (flags) + 32 (weights) read, 32 (ms) + 16
    (weights) written; epilogue 32 read + 64 written; uvw 24 B per row. CPU baseline: the
    oracle's numpy pieces (phasor, mask, polarisation matrix) on the visibilities of 65,536
    rows."""
    import ctypes

    import numpy
    import torch

    from rascil_amd import _lib, simulation
    from rascil_amd.data_models import Direction, PolarisationFrame, direction_to_lmn
    from rascil_amd.ng import _conv

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    lib = _lib.load()
    uvw, freq, pc = simulation.config1_geometry()
    nrow, nchan, npol = uvw.shape[0], len(freq), 4
    nvis = nrow * nchan * npol
    l, m, nm1 = direction_to_lmn(Direction(pc.ra + math.radians(0.5) / math.cos(pc.dec), pc.dec), pc)
    lmn = numpy.array([l, m, nm1])
    conv_in, _ = _conv(PolarisationFrame("linear"), PolarisationFrame("stokesIQUV"))
    conv_out, _ = _conv(PolarisationFrame("stokesIQUV"), PolarisationFrame("linear"))
    fq = numpy.ascontiguousarray(freq, "f8")
    pd = ctypes.POINTER(ctypes.c_double)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    uvw_d = torch.from_numpy(numpy.ascontiguousarray(uvw)).to(dev)
    out = {"study": "vis_passes", "nvis": int(nvis), "rows": int(nrow), "channels": nchan, "pols": npol,
           "frames": "linear visibilities, stokesIQUV image", "shift_deg": 0.5, "gridder_precision": "mixed",
           "data": "synthetic"}

    def timed(fn, name, nbytes):
        fn()
        torch.cuda.synchronize()
        _lib.profile_enable(True)
        ms = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        prof = _lib.profile_read()
        _lib.profile_enable(False)
        t = float(numpy.median(ms))
        res = {"ms_per_call": round(t, 3), "value": round(nvis / (t * 1e-3) / 1e6, 1), "unit": "Mvis/s",
               "kernel_ms": {k: round(tot / cnt, 3) for k, (tot, cnt) in prof.items()},
               "roofline": {"bound": "hbm", "achieved": round(nbytes / (t * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(nbytes / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "algorithmic_bytes_per_call": int(nbytes)}}
        out[name] = res
        print("[study] vis_passes", name, res, file=sys.stderr, flush=True)

    n = nrow * nchan
    vis = torch.randn((nrow, nchan, npol), dtype=torch.complex128, device=dev, generator=g)
    flags = (torch.rand((nrow, nchan, npol), device=dev, generator=g) < 0.05).to(torch.int32)
    iwt = torch.rand((nrow, nchan, npol), dtype=torch.float64, device=dev, generator=g) + 0.5
    ms = torch.empty((nrow, nchan, npol), dtype=torch.complex64, device=dev)
    wgt = torch.empty((nrow, nchan, npol), dtype=torch.float32, device=dev)
    nz = torch.empty((nchan, npol), dtype=torch.int32, device=dev)

    def pro():
        _lib.check(lib.hipng_vis_prologue(
            ctypes.c_void_p(vis.data_ptr()), ctypes.c_void_p(flags.data_ptr()), ctypes.c_void_p(iwt.data_ptr()),
            ctypes.c_void_p(uvw_d.data_ptr()), fq.ctypes.data_as(pd), nrow, nchan, npol, npol,
            conv_in.ctypes.data_as(pd), lmn.ctypes.data_as(pd), _lib.PRECISIONS["mixed"],
            ctypes.c_void_p(ms.data_ptr()), ctypes.c_void_p(wgt.data_ptr()), ctypes.c_void_p(nz.data_ptr()), st))

    timed(pro, "prologue", 160.0 * n + 24.0 * nrow)
    vis_h = vis[:65536].cpu().numpy()
    flags_h = flags[:65536].cpu().numpy()
    del flags, iwt, wgt, nz
    torch.cuda.empty_cache()

    def epi():
        _lib.check(lib.hipng_vis_epilogue(
            ctypes.c_void_p(ms.data_ptr()), ctypes.c_void_p(uvw_d.data_ptr()), fq.ctypes.data_as(pd), nrow, nchan,
            npol, npol, conv_out.ctypes.data_as(pd), lmn.ctypes.data_as(pd), _lib.PRECISIONS["mixed"],
            ctypes.c_void_p(vis.data_ptr()), st))

    timed(epi, "epilogue", 96.0 * n + 24.0 * nrow)
    # CPU baseline: the oracle's pieces on 8,192 rows (prologue then epilogue)
    from oracle import explicit as ox

    rows = 65536
    ul = uvw[:rows, None, :] * (freq / 299792458.0)[None, :, None]
    t0 = time.time()
    ph = ox.phasor(ul, l, m, nm1)
    x = ox.convert_pol(vis_h * (1 - flags_h) * numpy.conj(ph)[..., None], "linear", "stokesIQUV", polaxis=-1)
    x.astype(numpy.complex64)
    ox.convert_pol(x.astype(numpy.complex64).astype(complex), "stokesIQUV", "linear", polaxis=-1) * ph[..., None]
    dt = time.time() - t0
    cpu = rows * nchan * npol / dt / 1e6
    out["cpu_baseline"] = {"value": round(cpu, 2), "unit": "Mvis/s", "cores": 1, "kind": "port",
                           "sample": f"numpy oracle pieces (prologue + epilogue) on {rows} rows x {nchan} ch x "
                                     f"{npol} pol ({dt:.2f} s)"}
    gpu = nvis / ((out["prologue"]["ms_per_call"] + out["epilogue"]["ms_per_call"]) * 1e-3) / 1e6
    out["prologue_plus_epilogue_mvis_per_s"] = round(gpu, 1)
    out["speedup_vs_cpu"] = round(gpu / cpu, 1)
    return out


def dft_study(reps=3, seed=19, ncomp=100, nrows=2 ** 21):
    """FP64 point-source DFT (hipng_dft): 2,097,152 rows of the production geometry (its first
    rows) x 64 channels x 4 polarisations, 100 components within 2 deg, complex128 fluxes;
    inputs resident in HBM. Per (visibility, component): the phase (3 FMA), one FP64 sincospi
    and 4 complex multiply-adds; the roofline counts 38 FLOP per pair (the sincospi not
    included) against the FP64 vector peak. CPU baseline: the oracle's numpy DFT on 8,192 rows
    with all components."""
    import ctypes

    import numpy
    import torch

    from rascil_amd import _lib, simulation

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    lib = _lib.load()
    uvw, freq, _ = simulation.config1_geometry()
    uvw = uvw[:nrows]
    nchan, npol = len(freq), 4
    rng = numpy.random.default_rng(seed)
    r = numpy.radians(2.0) * numpy.sqrt(rng.uniform(0, 1, ncomp))
    phi = rng.uniform(0, 2 * numpy.pi, ncomp)
    l, m = r * numpy.cos(phi), r * numpy.sin(phi)
    dc = numpy.stack([l, m, numpy.sqrt(1 - l * l - m * m) - 1], axis=1)
    flux = (rng.uniform(0.1, 1.0, (ncomp, nchan, npol)) + 0j)
    uvl = uvw[:, None, :] * (freq / 299792458.0)[None, :, None]
    dc_d = torch.from_numpy(dc).to(dev)
    fl_d = torch.from_numpy(flux).to(dev)
    uvl_d = torch.from_numpy(numpy.ascontiguousarray(uvl)).to(dev)
    vis = torch.empty((nrows, nchan, npol), dtype=torch.complex128, device=dev)
    st = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    nvis = nrows * nchan * npol

    def call():
        _lib.check(lib.hipng_dft(ctypes.c_void_p(dc_d.data_ptr()), ctypes.c_void_p(fl_d.data_ptr()),
                                 ctypes.c_void_p(uvl_d.data_ptr()), ncomp, nrows, nchan, npol,
                                 ctypes.c_void_p(vis.data_ptr()), st))

    call()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        call()
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    t = float(numpy.median(ms)) * 1e-3
    flop = 38.0 * nrows * nchan * ncomp
    out = {"study": "dft", "nvis": int(nvis), "rows": nrows, "channels": nchan, "pols": npol,
           "components": ncomp, "data": "synthetic", "ms_per_call": round(t * 1e3, 2),
           "value": round(nvis / t / 1e6, 1), "unit": "Mvis/s",
           "vis_components_per_s": round(nrows * nchan * ncomp / t / 1e9, 2),
           "roofline": {"bound": "fp64", "achieved": round(flop / t / 1e12, 2), "peak": FP64_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": round(flop / t / 1e12 / FP64_PEAK_TFLOPS, 4),
                        "flop_per_call": int(flop), "note": "38 FLOP per (visibility row-channel, component); "
                                                            "sincospi not counted"}}
    print("[study] dft", out, file=sys.stderr, flush=True)
    from oracle import explicit as ox

    rows = 8192
    t0 = time.time()
    ox.dft_cpu_looped(dc, uvl[:rows][None], flux)
    dt = time.time() - t0
    cpu = rows * nchan * npol / dt / 1e6
    out["cpu_baseline"] = {"value": round(cpu, 3), "unit": "Mvis/s", "cores": 1, "kind": "port",
                           "sample": f"numpy oracle DFT on {rows} rows x {nchan} ch x {npol} pol x {ncomp} "
                                     f"components ({dt:.2f} s)"}
    out["speedup_vs_cpu"] = round(out["value"] / cpu, 1)
    return out


def copy_probe(nbytes=1 << 30, reps=4):
    """A plain device copy of known size: the rocprof FETCH_SIZE / WRITE_SIZE counters read on
    it give the scale factor between counter units and bytes moved."""
    import torch

    dev = torch.device("cuda", 0)
    a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t = time.time()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.time() - t) / reps
    return {"probe": "copy", "bytes_read": nbytes, "bytes_written": nbytes, "reps": reps,
            "GBps": round(2 * nbytes / dt / 1e9, 1)}


def main():
    args = parse()
    if args.copy_probe:
        print(json.dumps(copy_probe()), flush=True)
        return
    if args.study:
        fn = {"precision": precision_study, "autograd": autograd_study, "batched": batched_study,
              "weighting": weighting_study, "gains": gains_study, "vis_passes": vis_passes_study,
              "dft": dft_study}[args.study]
        print(json.dumps(fn()), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus %d is not supported "
                                   "(invert shards by channel or row block with one all-reduce of the image; predict by rows "
                                   "after one broadcast of the model)" % args.gpus}))
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
    opts = {k: v for k, v in (("sigma", args.sigma), ("tile", args.tile)) if v} or None
    print(f"[bench] {wl['name']}: {nrow} rows x {nchan} ch x {npol} pol = {nvis:.3e} vis, "
          f"{N}^2, precision {prec}", file=sys.stderr, flush=True)
    cdt = torch.complex128 if prec == "fp64" else torch.complex64
    rdt = torch.float64 if prec == "fp64" else torch.float32
    g = torch.Generator(device=dev)
    g.manual_seed(1234)
    uvw_d = torch.from_numpy(numpy.ascontiguousarray(wl["uvw"])).to(dev)
    vis_d = torch.randn((nrow, nchan, npol), dtype=cdt, device=dev, generator=g)
    # imaging weights for the invert, as RASCIL passes them (ng.py:232-248); predict has none
    wgt_d = torch.empty((nrow, nchan, npol), dtype=rdt, device=dev).uniform_(0.5, 1.5, generator=g)
    model_d = torch.randn((npol, N, N), dtype=rdt, device=dev, generator=g)
    torch.cuda.synchronize()
    plan_i, plan_p = {}, {}

    def step(p, pi=None, pp=None):
        wgridder.ms2dirty(uvw_d, wl["freq"], vis_d, wgt_d, N, N, wl["cell"], wl["cell"], 0, 0,
                          wl["eps"], True, precision=p, plan_out=pi, plan_opts=opts)
        wgridder.dirty2ms(uvw_d, wl["freq"], model_d, None, wl["cell"], wl["cell"], 0, 0,
                          wl["eps"], True, precision=p, plan_out=pp, plan_opts=opts)

    for i in range(max(1, args.warmup)):
        t = time.time()
        last = i == max(1, args.warmup) - 1
        # the plans are read back on the last warm-up step only (that costs a device sync)
        step(prec, plan_i if last else None, plan_p if last else None)
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
    fp_peak = FP64_PEAK_TFLOPS if prec == "fp64" else FP32_PEAK_TFLOPS

    def plan_of(name):
        return plan_p if name in ("degrid", "degrid_dense", "screen_fill", "fft_c2c_forward") else plan_i

    kern = sorted(((tot, name, cnt) for name, (tot, cnt) in prof.items()), reverse=True)
    # the roofline object is the dominant kernel class written here (rocFFT, a library, is
    # reported beside it in `library_fft`)
    roof = None
    for tot, name, cnt in kern:
        if name.startswith("fft"):
            continue
        plan = plan_of(name)
        per_step = cnt / args.steps
        pc = max(1, round(plan["nplanes"] / per_step)) if name in ("grid", "grid_dense") else 1  # planes per pass
        flop, nbytes = algorithmic_work(name, plan, nvis, nrow * nchan, npol, csz, pc, "degrid_dense" in prof)
        if flop is None and nbytes is None:
            continue
        avg_s = tot / cnt * 1e-3
        traffic, src = measured_traffic(name, plan, prec, npol)
        alg_bytes = nbytes / per_step
        hbm_view = {"achieved": round(alg_bytes / avg_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "algorithmic_bytes_per_launch": int(alg_bytes)}
        hbm_view["frac"] = round(hbm_view["achieved"] / HBM_PEAK_GBS, 4)
        if traffic:
            hbm_view["traffic_over_algorithmic"] = round(traffic / alg_bytes, 2)
        if flop is not None:
            # FP32 modes accumulate on the matrix cores (v_mfma_f32_16x16x4_f32, dense FP32
            # matrix peak); FP64 gridding is vector FMA (dense FP64 peak)
            ach = flop / per_step / avg_s / 1e12
            roof = {"kernel": name, "bound": "mfma", "achieved": round(ach, 2), "peak": fp_peak,
                    "unit": "TFLOP/s", "frac": round(ach / fp_peak, 4), "traffic": traffic,
                    "avg_launch_ms": round(tot / cnt, 4), "launches": cnt,
                    "algorithmic_flop_per_launch": int(flop / per_step),
                    "compute": ("vector FP64 FMA; peak = dense FP64" if prec == "fp64" else
                                "FP32 MFMA (v_mfma_f32_16x16x4_f32) plus vector tap evaluation; "
                                "peak = dense FP32 matrix"),
                    "hbm_view": hbm_view}
        else:
            roof = {"kernel": name, "bound": "hbm", "achieved": hbm_view["achieved"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": hbm_view["frac"], "traffic": traffic,
                    "avg_launch_ms": round(tot / cnt, 4), "launches": cnt,
                    "algorithmic_bytes_per_launch": hbm_view["algorithmic_bytes_per_launch"]}
        if src:
            roof["traffic_source"] = "profiles/" + src
        break
    breakdown = {name: {"total_ms": round(tot, 2), "launches": cnt} for tot, name, cnt in kern}
    lib_fft = {}
    for tot, name, cnt in kern:
        if not name.startswith("fft"):
            continue
        plan = plan_of(name)
        _, nbytes = algorithmic_work(name, plan, nvis, nrow * nchan, npol, csz, 1)
        gbs = nbytes * args.steps / (tot * 1e-3) / 1e9
        lib_fft[name] = {"ms_per_step": round(tot / args.steps, 2), "launches_per_step": cnt // args.steps,
                         "bytes_per_step": int(nbytes), "achieved_gbs": round(gbs, 1),
                         "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mvis/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32", "fp32": "f32"}[prec], "data": "synthetic",
        "config": {"workload": wl["desc"], "nvis": nvis, "precision": prec, "invert_weights": True,
                   "W": plan_i.get("W"), "sigma": plan_i.get("sigma"), "grid": plan_i.get("gx"),
                   "nplanes": plan_i.get("nplanes"), "tile": plan_i.get("tile"),
                   "dense_min": plan_i.get("dense_min"),
                   "dense_fraction": round(plan_i.get("dense_items", 0) / max(1, nrow * nchan), 4)},
        "roofline": roof,
        "library_fft": lib_fft,
        "path_roofline": path_roofline(plan_i, plan_p, nrow, nchan, npol, prec, True, ms_per_step),
        "kernel_ms_total": breakdown,
    }

    if not args.no_accuracy:
        try:
            out["accuracy"] = accuracy_at_timed_plan(wl, prec, uvw_d, vis_d, wgt_d, (plan_i, plan_p))
            print(f"[bench] accuracy {out['accuracy']}", file=sys.stderr, flush=True)
        except Exception as e:
            out["accuracy"] = {"error": repr(e)}

    if args.config == "c1" and args.rascil_steps > 0:
        # the same workload through the RASCIL-level API on device-resident columns
        del vis_d, wgt_d
        torch.cuda.empty_cache()
        _lib.load().hipng_release()
        try:
            out["rascil_level"] = rascil_level_leg(wl, prec, uvw_d, args.rascil_steps, ms_per_step)
            print(f"[bench] rascil level {out['rascil_level']}", file=sys.stderr, flush=True)
        except Exception as e:
            out["rascil_level"] = {"error": repr(e)}
        torch.cuda.empty_cache()
        vis_d = wgt_d = None

    if args.fp64_steps > 0 and prec != "fp64":
        vis_d = model_d = wgt_d = None
        torch.cuda.empty_cache()
        _lib.load().hipng_release()
        # FP64 buffers of all four polarisations at once do not fit next to the FP64 records in
        # HBM: two jobs of two polarisations each (the same work as one four-polarisation call)
        ngrp = 2 if nvis > 1e9 else 1
        gp = npol // ngrp
        vis64 = [torch.randn((nrow, nchan, gp), dtype=torch.complex128, device=dev, generator=g)
                 for _ in range(ngrp)]
        wgt64 = torch.empty((nrow, nchan, gp), dtype=torch.float64, device=dev).uniform_(0.5, 1.5, generator=g)
        mod64 = [torch.randn((gp, N, N), dtype=torch.float64, device=dev, generator=g)
                 for _ in range(ngrp)]
        p64 = ({}, {})

        def step64(pi=None, pp=None):
            for v64, m64 in zip(vis64, mod64):
                wgridder.ms2dirty(uvw_d, wl["freq"], v64, wgt64, N, N, wl["cell"], wl["cell"], 0, 0,
                                  wl["eps"], True, precision="fp64", plan_out=pi)
                wgridder.dirty2ms(uvw_d, wl["freq"], m64, None, wl["cell"], wl["cell"], 0, 0,
                                  wl["eps"], True, precision="fp64", plan_out=pp)

        step64(*p64)
        torch.cuda.synchronize()
        t = time.time()
        for i in range(args.fp64_steps):
            step64()
            print(f"[bench] fp64 step {i} done", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        ms64 = (time.time() - t) * 1e3 / args.fp64_steps
        out["fp64_plain"] = {"ms_per_step": round(ms64, 1), "value": round(nvis / (ms64 * 1e-3) / 1e6, 2),
                             "speedup_of_fp32_mode": round(ms64 / ms_per_step, 2),
                             "jobs": ngrp, "pols_per_job": gp, "W": p64[0].get("W"), "sigma": p64[0].get("sigma"),
                             "grid": p64[0].get("gx"), "nplanes": p64[0].get("nplanes")}
        if not args.no_accuracy:
            try:
                del vis64[1:], mod64
                torch.cuda.empty_cache()
                out["fp64_plain"]["accuracy"] = accuracy_at_timed_plan(wl, "fp64", uvw_d, vis64[0], wgt64, p64)
                print(f"[bench] fp64 accuracy {out['fp64_plain']['accuracy']}", file=sys.stderr, flush=True)
            except Exception as e:
                out["fp64_plain"]["accuracy"] = {"error": repr(e)}
        del vis64, wgt64
        torch.cuda.empty_cache()
        _lib.load().hipng_release()

    if not args.no_cpu:
        try:
            from oracle import wstack_cpu

            # the small configuration runs in full; the production shape on 1 % of its rows
            out["cpu_baseline"] = wstack_cpu.cpu_baseline(wl, plan_i, plan_p, args.cpu_seconds,
                                                          plans=cpu_plans(plan_i, npol, nrow * nchan),
                                                          frac=1.0 if nvis < 1e7 else 0.01)
        except Exception as e:  # the baseline leg must not hide the GPU result
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
