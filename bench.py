#!/usr/bin/env python3
"""Benchmark of the model-image hot path on one MI355X.

Workload (one "step"): a batch of B likelihood+gradient evaluations of the
400x400 bulge+disk galaxy (Sersic n=4, re=3 + Sersic n=1, re=15) convolved
with a 51x51 Moffat PSF (FWHM 4, con 3), chi^2 likelihood over an elliptical
region of ~70% of the pixels, gradient w.r.t. the 8 parameters of each
component.  The B parameter sets are walkers scattered around the truth.

Printed JSON line: value = batched-adjoint lik+grad evaluations/s; beside it
the plain per-call GPU forward path with central finite differences (29
forward likelihoods per lik+grad), the parity of the GPU result against the
FP64 oracle, the roofline of the dominant kernel from HIP-event timings on the
engine stream over the timed steps, and the FP64 CPU oracle (libprofit's CPU
algorithm, OpenMP on the host cores) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "lik+grad evals/s, 400² bulge+disk ⊛ PSF: adjoint vs FD, ≤1e-5 rel pixel err"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def walkers(base, n, seed=11):
    """n parameter sets scattered around ``base`` (list of component dicts)."""
    from profit_amd import workloads as W
    u = (W.splitmix64(seed, n * len(base) * 7) >> np.uint64(11)).astype(np.float64) / 9007199254740992.0
    u = (u - 0.5).reshape(n, len(base), 7)
    out = []
    for i in range(n):
        comps = []
        for k, c in enumerate(base):
            d = dict(c)
            d["xcen"] += 0.4 * u[i, k, 0]
            d["ycen"] += 0.4 * u[i, k, 1]
            d["mag"] += 0.1 * u[i, k, 2]
            d["re"] *= 1 + 0.04 * u[i, k, 3]
            d["nser"] *= 1 + 0.04 * u[i, k, 4]
            d["ang"] += 4.0 * u[i, k, 5]
            d["axrat"] = min(1.0, d["axrat"] * (1 + 0.04 * u[i, k, 6]))
            comps.append(d)
        out.append(comps)
    return out


def gpu_params(P, sets):
    flat = []
    for comps in sets:
        for c in comps:
            c = dict(c)
            t = c.pop("type")
            flat.append(P.make_profile(t, convolve=True, **c))
    return P.profile_array(flat)


def fd_gradient_gpu(eng, P, comps, slots=("xcen", "ycen", "mag", "re", "nser", "ang", "axrat"), rel=1e-4):
    """Central finite differences through the plain forward path: 1 + 2*14 likelihoods."""
    ll0 = eng.like(gpu_params(P, [comps]))[0]
    grads = []
    for k in range(len(comps)):
        for s in slots:
            h = rel * abs(comps[k][s]) if comps[k][s] != 0 else rel
            vals = []
            for sign in (1, -1):
                cc = [dict(c) for c in comps]
                cc[k][s] += sign * h
                vals.append(eng.like(gpu_params(P, [cc]))[0])
            grads.append((vals[0] - vals[1]) / (2 * h))
    return ll0, np.array(grads)


def cpu_baseline(wl, calc, seconds, threads):
    """FP64 oracle: lik+grad by central differences (29 likelihoods), timed on
    the host cores; plus its model, logL and analytic gradient for parity."""
    import oracle as O
    from tests.helpers import oracle_profiles
    spec = O.ModelSpec(400, 400, psf=wl["psf"], mask=calc, adjust_mask=False, magzero=wl["magzero"],
                       threads=threads)
    fit = O.FitSpec(spec, wl["data"], wl["sigma"], wl["region"], "chisq")
    comps = wl["components"]
    slots = np.zeros((len(comps), O.NGRAD), dtype=np.uint8)
    slots[:, [0, 1, 2, 3, 4, 6, 7]] = 1
    n = 0
    t0 = time.perf_counter()
    while True:
        O.loglike_grad_fd(fit, oracle_profiles(comps), rel_step=1e-4, slots=slots, frozen=False)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 3:
            break
    ll, model = O.loglike(fit, oracle_profiles(comps), want_model=True)
    _, grad = O.loglike_grad(fit, oracle_profiles(comps))
    return {"value": n / el, "elapsed": el, "n": n, "ll": ll, "model": model, "grad": grad}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--fd-evals", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=8)
    ap.add_argument("--render-fp32", action="store_true", help="direct pixels in FP32 instead of FP64")
    ap.add_argument("--conv", choices=("auto", "brute", "fft", "brute-bf16x3"), default="auto",
                    help="PSF convolution method (auto: the autotuner's choice among brute and fft)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import profit_amd as P
    from profit_amd import workloads as W

    def gpu_model_fn(w, h, psf, comps, magzero):
        return P.make_model(w, h, gpu_params(P, [comps]), psf=psf, magzero=magzero)["image"]

    wl = W.config2(gpu_model_fn, magzero=30.0)
    calc = P.adjust_mask(wl["region"], (400, 400), wl["psf"])
    B = args.batch
    eng = P.Engine()
    g = eng.add_galaxy(400, 400, wl["data"], wl["sigma"], psf=wl["psf"], calcregion=calc, region=wl["region"],
                       likefunc="chisq", magzero=30.0)
    eng.set_tile_rows(args.tile_rows)
    eng.set_render_fp64(not args.render_fp32)
    eng.set_conv_method("brute" if args.conv == "auto" else args.conv)
    eng.set_batch([g] * B, [2] * B)
    sets = walkers(wl["components"], B)
    params = gpu_params(P, sets)
    tune = None
    if args.conv == "auto":
        tune = eng.autotune(params, reps=2)
        print(f"[bench] autotune: {tune['kinds'][0]} ({tune['ms']})", file=sys.stderr, flush=True)
    conv_used = tune["kinds"][0] if tune else args.conv
    print(f"[bench] batch {B}: warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        eng.like_grad(params)

    # timed region: eager launch sequence with HIP events on the engine stream
    eng.set_timing(True)
    eng.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ll, grad = eng.like_grad(params)
    t1 = time.perf_counter()
    stats = eng.stats()
    eng.set_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    print(f"[bench] adjoint: {value:.1f} lik+grad/s, {ms_per_step:.2f} ms/step", file=sys.stderr, flush=True)

    # graph replay of the same batch (production path)
    eng.like_grad(params)
    tg0 = time.perf_counter()
    for _ in range(args.steps):
        eng.like_grad(params)
    tg1 = time.perf_counter()
    graph_value = B * args.steps / (tg1 - tg0)

    # plain per-call GPU forward path + central finite differences
    eng1 = P.Engine()
    g1 = eng1.add_galaxy(400, 400, wl["data"], wl["sigma"], psf=wl["psf"], calcregion=calc, region=wl["region"],
                         likefunc="chisq", magzero=30.0)
    eng1.set_render_fp64(not args.render_fp32)
    eng1.set_conv_method(conv_used)
    eng1.set_batch([g1], [2])
    fd_gradient_gpu(eng1, P, sets[0])  # warm-up (graph capture)
    tf0 = time.perf_counter()
    for i in range(args.fd_evals):
        fd_gradient_gpu(eng1, P, sets[i % B])
    tf1 = time.perf_counter()
    fd_value = args.fd_evals / (tf1 - tf0)
    print(f"[bench] plain forward + FD: {fd_value:.1f} lik+grad/s", file=sys.stderr, flush=True)

    # roofline of the dominant kernel (per launch = one batch)
    sm = stats["stage_ms"]
    calls = max(1, stats["stage_calls"]["conv_like"])
    fwd_ms = sm["conv_like"] / calls
    adj_ms = sm["adjoint_conv"] / max(1, stats["stage_calls"]["adjoint_conv"])
    if adj_ms >= fwd_ms:
        kname, flops, kms = "k_conv<8,CONV_ADJ> (adjoint correlation)", stats["adjoint_flops"], adj_ms
    else:
        kname, flops, kms = "k_conv<8,CONV_LIKE> (convolution + likelihood)", stats["conv_flops"], fwd_ms
    achieved = flops / (kms * 1e-3) / 1e12
    stage_per_step = {k: v / args.steps for k, v in sm.items()}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "lik+grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (render, convolution) + f64 (classification, recursion, likelihood sums)",
        "data": "synthetic (noiseless model from this GPU path plus seeded SplitMix64/Box-Muller noise)",
        "config": {
            "workload": "400x400 bulge+disk (Sersic n=4 re=3 mag=17 axrat=0.9; Sersic n=1 re=15 mag=16 "
                        "axrat=0.4 ang=60) convolved with a 51x51 Moffat PSF (FWHM 4, con 3); chi^2 over an "
                        "elliptical region (70% of pixels); magzero 30; adjoint gradient of 8 parameters per "
                        "component",
            "batch": B,
            "drawing_canvas": "450x450",
            "tile_rows": args.tile_rows,
            "direct_pixels": "fp32" if args.render_fp32 else "fp64",
            "convolution": conv_used,
            "autotune_ms": tune["ms"] if tune else None,
        },
        "adjoint_vs_fd": {
            "adjoint_graph_replay_value": round(graph_value, 3),
            "fd_plain_forward_value": round(fd_value, 3),
            "ratio": round(value / fd_value, 2),
            "fd_forward_calls_per_lik_grad": 29,
        },
        "stage_ms_per_step": {k: round(v, 4) for k, v in stage_per_step.items()},
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": None,
            "kernel": kname,
            "flops_per_launch": flops,
            "avg_launch_ms": round(kms, 4),
            "note": "FP32 FMA on the vector ALUs (packed FP32 disabled); peak quoted is MI355X FP32 157.3 TF",
        },
    }

    # FP64 CPU oracle leg: baseline timing and parity of the GPU result
    if not args.no_cpu_baseline:
        threads = os.cpu_count() or 1
        cb = cpu_baseline(wl, calc, args.cpu_seconds, threads)
        eng1.set_batch([g1], [2])
        truth = gpu_params(P, [wl["components"]])
        gmodel = eng1.model_images(truth)[0]
        gll, ggrad = eng1.like_grad(truth)
        ref = cb["model"]
        peak = np.abs(ref).max()
        sel = np.abs(ref) > 1e-6 * peak
        pix_err = float(np.max(np.abs(gmodel[sel] - ref[sel]) / np.abs(ref[sel])))
        gslots = [0, 1, 2, 3, 4, 6, 7]
        og = cb["grad"][:, gslots]
        gg = ggrad[:, gslots]
        grad_err = float(np.max(np.abs(gg - og) / np.maximum(np.abs(og), 1e-3 * np.abs(og).max())))
        out["parity_vs_fp64_oracle"] = {
            "max_rel_pixel_err": pix_err,
            "pixels_checked": int(sel.sum()),
            "abs_logL_err": float(abs(gll[0] - cb["ll"])),
            "max_rel_grad_err": grad_err,
            "gates": {"pixel": 1e-5, "logL": 1e-3},
        }
        out["cpu_baseline"] = {
            "value": round(cb["value"], 5),
            "unit": "lik+grad evals/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{cb['n']} lik+grad evaluation(s) of the same workload by central differences "
                      f"(29 FP64 likelihoods each) in {cb['elapsed']:.1f} s, OpenMP over pixels",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
