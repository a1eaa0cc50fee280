#!/usr/bin/env python3
"""Benchmark of the layer optical-depth line sum on one MI355X.

Workload (the full-profile OD case): 60 layers with log-spaced pressure from 1013
to 0.1 hPa, US-standard-like temperatures, 7 species (HITRAN molecules 1-7),
200 000 synthetic lines over 500-2500 cm-1, per-layer DV from calculateDV
(mean Voigt half-width / 4), a 50-knot spline continuum.  One step = the whole
layer-OD pass over all 60 layers (one batched set of launches), inputs resident
in HBM.

Metric: Voigt line-evals/s, E / step time, where
E = sum over layers and lines of the number of layer-grid points within
+-25 cm-1 of the shifted line centre (a closed form of the line list and grids).

Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Voigt line-evals/s & Jacobian spectra/s vs FP64 GPU path at ≤1e-5 OD err"
V1, V2 = 500.0, 2500.0
# MI355X datasheet peaks (vector FP64 and FP32 without packed math are the same issue rate)
PEAK = {"fp64": 78.6, "fp32": 157.3, "hbm": 8000.0}
FLOP_CONV = 13   # per R1/R2/R3 evaluation: 3 table-argument adds, 5 products, 5 sums
FLOP_F4 = 7      # per (line, R4 point): offset, square, scaled square, 3 for the shape, 1 sum


def workload():
    from clblm_amd import workloads as W
    lines = W.profile_lines(200000, V1, V2, seed=1001)
    layers = W.atmosphere(60, 0.1)
    vbar = 0.5 * (V1 + V2)
    dvs = [W.calculate_dv(vbar, l.pave, l.tave, l.w[1] / l.wtot) for l in layers]
    conts = [W.continuum_absrb(l, V1, V2)[0] for l in layers]
    return lines, layers, dvs, conts


def line_evals(lines, layers, dvs):
    """E = sum_l sum_j #{i : |v1 + i dv_l - nu_j| <= 25, 0 <= i < n_l}."""
    nu = lines.nu
    tot = 0
    per = []
    for l, dv in zip(layers, dvs):
        n_pts = int(math.ceil((V2 - V1) / dv)) + 2
        lo = np.ceil((nu - 25.0 - V1) / dv)
        hi = np.floor((nu + 25.0 - V1) / dv)
        lo = np.clip(lo, 0, n_pts - 1)
        hi = np.clip(hi, -1, n_pts - 1)
        c = int(np.maximum(hi - lo + 1, 0).sum())
        per.append(c)
        tot += c
    return tot, per


def cpu_baseline(lines, layers, dvs, conts, per_layer_evals, budget_s=20.0):
    """Oracle (sequential C restatement, one core) on a bounded sample of layers."""
    from oracle.py import Oracle
    order = list(range(0, len(layers), 10)) + list(range(5, len(layers), 10))
    t_tot, e_tot, used = 0.0, 0, []
    orc = Oracle(lines)
    for k in order:
        t0 = time.perf_counter()
        orc.odlay(layers[k], V1, V2, dvs[k], continuum=conts[k])
        t_tot += time.perf_counter() - t0
        e_tot += per_layer_evals[k]
        used.append(k)
        if t_tot > budget_s:
            break
    return e_tot / t_tot, used, t_tot


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32c"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    from clblm_amd.odlay import Context
    from clblm_amd._lib import FP32C, FP64

    lines, layers, dvs, conts = workload()
    E, per = line_evals(lines, layers, dvs)
    prec = FP64 if args.precision == "fp64" else FP32C
    ctx = Context(0)
    h = ctx.upload_lines(lines)
    pars = [Context.params(V1, V2, dv, precision=prec, continuum=c) for dv, c in zip(dvs, conts)]
    ctx.jrad = 1
    out, off = ctx.od_batch(h, layers, pars)
    for _ in range(max(0, args.warmup - 1)):
        ctx.jrad = 1
        ctx.od_batch(h, layers, pars, out=out)
    torch.cuda.synchronize()
    ctx.work(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.jrad = 1
        ctx.od_batch(h, layers, pars, out=out)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    w = ctx.work(reset=False)
    tm = ctx.timing(reset=True)
    value = E / dt
    conv_tf = w["conv_evals"] * FLOP_CONV / (w["ms_conv"] * 1e-3) / 1e12 if w["ms_conv"] > 0 else 0.0
    f4_tf = w["f4_evals"] * FLOP_F4 / (w["ms_f4"] * 1e-3) / 1e12 if w["ms_f4"] > 0 else 0.0
    dom = "k_f4conv" if w["ms_f4"] >= w["ms_conv"] else "k_conv"
    ach = f4_tf if dom == "k_f4conv" else conv_tf
    peak = PEAK["fp64"] if prec == FP64 else PEAK["fp32"]
    res = {
        "metric": METRIC, "value": value, "unit": "line-evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == FP64 else "f32+f64",
        "data": "synthetic (seeded line list and atmosphere; no TAPE3/scene in the reference)",
        "config": {"workload": "60-layer full-profile OD: 200k lines of 7 species over 500-2500 cm-1, "
                               "1013-0.1 hPa, DV = mean Voigt HWHM/4, spline continuum",
                   "layers": len(layers), "lines": lines.n, "points": int(off[-1]),
                   "precision": args.precision, "line_evals_per_step": E},
        "roofline": {"bound": "valu", "kernel": dom, "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                     "frac": ach / peak, "traffic": None,
                     "evals": {"conv": w["conv_evals"], "f4": w["f4_evals"]},
                     "ms": {"conv_kernel": w["ms_conv"], "f4_kernel": w["ms_f4"], "prep": tm["prep"],
                            "conv_stage": tm["conv"], "f4_stage": tm["f4"], "panel_stage": tm["panel"]}},
    }
    if not args.no_cpu:
        v, used, t = cpu_baseline(lines, layers, dvs, conts, per)
        res["cpu_baseline"] = {"value": v, "unit": "line-evals/s", "cores": 1, "kind": "port",
                               "sample": "oracle (sequential C restatement of ODLAY, -O2) on layers %s of the same "
                                         "workload, full 500-2500 cm-1 range, %.1f s" % (used, t)}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
