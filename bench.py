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
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
# the 14-parameter gradient of the bulge+disk metric: 7 slots per component
METRIC_SLOTS = ("xcen", "ycen", "mag", "ang", "axrat", "s0", "s1")


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
        if el >= seconds:
            break
    ll, model = O.loglike(fit, oracle_profiles(comps), want_model=True)
    _, grad = O.loglike_grad(fit, oracle_profiles(comps))
    return {"value": n / el, "elapsed": el, "n": n, "ll": ll, "model": model, "grad": grad}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--fd-evals", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=8)
    ap.add_argument("--render-fp64", action="store_true",
                    help="profile values (direct pixels, recursion leaves) in FP64 instead of FP32")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 256 MiB device-to-device three times (the reference kernel for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--conv", choices=("auto", "brute", "fft", "brute-bf16x3"), default="auto",
                    help="PSF convolution method (auto: the autotuner's choice among brute and fft)")
    ap.add_argument("--grad-slots", choices=("metric", "all"), default="metric",
                    help="gradient computed per component: 'metric' = the 7 parameters of the 14-parameter "
                         "bulge+disk gradient (xcen, ycen, mag, re, nser, ang, axrat); 'all' adds box and "
                         "the unused shape slots")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = 64 * 1024 * 1024
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 4}))
        return
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
    eng.set_render_fp64(args.render_fp64)
    eng.set_conv_method("brute" if args.conv == "auto" else args.conv)
    slots = METRIC_SLOTS if args.grad_slots == "metric" else None
    eng.set_grad_slots(slots)
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
    eng1.set_render_fp64(args.render_fp64)
    eng1.set_conv_method(conv_used)
    eng1.set_grad_slots(slots)
    eng1.set_batch([g1], [2])
    fd_value = None
    if args.fd_evals > 0:
        fd_gradient_gpu(eng1, P, sets[0])  # warm-up (graph capture)
        tf0 = time.perf_counter()
        for i in range(args.fd_evals):
            fd_gradient_gpu(eng1, P, sets[i % B])
        tf1 = time.perf_counter()
        fd_value = args.fd_evals / (tf1 - tf0)
    if fd_value:
        print(f"[bench] plain forward + FD: {fd_value:.1f} lik+grad/s", file=sys.stderr, flush=True)

    # roofline of the dominant kernel (one launch = the whole batch), with the
    # algorithmic work per unit of each kernel family:
    #   rendering (k_render, the recursion walkers, the adjoint render):
    #     5 B per canvas pixel (4 B FP32 write + 1 B mask read), HBM-bound
    #   brute-force convolution: 2 K^2 FLOP per output pixel (FP32 vector peak)
    #   FFT convolution (3 kernels): 3 transforms x 5 N log2 N FLOP, N = N1 N2
    #     (FP64 vector peak)
    sm = stats["stage_ms"]
    ncalls = max(1, stats["stage_calls"]["subsample"])
    stage_ms = {k: v / ncalls for k, v in sm.items()}
    dw, dh, _, _ = eng.geometry(g)
    K = wl["psf"].shape[0]
    n1 = 1 << int(np.ceil(np.log2(dw + K)))
    n2 = 1 << int(np.ceil(np.log2(dh + K)))
    n_fft = n1 * n2
    render_bytes = B * dw * dh * 5.0
    kernels = {
        "render": ("k_render<0> (classification + direct pixels, Sersic)", "hbm", render_bytes, PEAK_HBM_GBS,
                   "GB/s"),
        "subsample": ("k_subsample<false, 0, 0> (recursion walker, Sersic)", "hbm", render_bytes, PEAK_HBM_GBS,
                      "GB/s"),
        "adjoint_render": ("k_adj_render<0, *> (adjoint of direct pixels, Sersic)", "hbm", render_bytes,
                           PEAK_HBM_GBS, "GB/s"),
        "adjoint_subsample": ("k_subsample<true, 0, 2> (adjoint recursion walker, Sersic)", "hbm", render_bytes,
                              PEAK_HBM_GBS, "GB/s"),
    }
    if conv_used == "fft":
        fft_flops = B * 3 * 5.0 * n_fft * np.log2(n_fft)
        kernels["conv_like"] = ("FFT convolution + likelihood (k_fft_rows_fwd, k_fft_cols, k_fft_rows_inv<0>)",
                                "mfma", fft_flops, PEAK_FP64_TFLOPS, "TFLOP/s")
        kernels["adjoint_conv"] = ("FFT adjoint correlation (k_fft_rows_fwd, k_fft_cols, k_fft_rows_inv<2>)", "mfma",
                                   fft_flops, PEAK_FP64_TFLOPS, "TFLOP/s")
    else:
        peak = PEAK_BF16_TFLOPS if conv_used == "brute-bf16x3" else PEAK_FP32_TFLOPS
        kname = "k_conv_mfma" if conv_used == "brute-bf16x3" else "k_conv"
        kernels["conv_like"] = (f"{kname}<CONV_LIKE> (convolution + likelihood)", "mfma", stats["conv_flops"], peak,
                                "TFLOP/s")
        kernels["adjoint_conv"] = (f"{kname}<CONV_ADJ> (adjoint correlation)", "mfma", stats["adjoint_flops"], peak,
                                   "TFLOP/s")
    dom = max(kernels, key=lambda k: stage_ms[k])
    kname, bound, work, peak, unit = kernels[dom]
    # HBM bytes per launch of that kernel from the committed PMC summary of the
    # same workload (FETCH_SIZE / WRITE_SIZE passes scaled by a copy kernel)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_config2_batch{B}.json")
    if os.path.exists(tfile) and conv_used == "fft":
        with open(tfile) as f:
            tk = json.load(f)["kernels"]
        key = kname.split(" (")[0]
        if key in tk:
            traffic = tk[key]["hbm_bytes_per_launch"]
        elif key.startswith("FFT"):
            # the stage's three kernels: rows forward, columns, rows inverse
            # (likelihood epilogue k_fft_rows_inv<0>, adjoint image <2>)
            inv = "k_fft_rows_inv<0>" if dom == "conv_like" else "k_fft_rows_inv<2>"
            parts = [k for k in ("k_fft_rows_fwd", "k_fft_cols", inv) if k in tk]
            traffic = sum(tk[k]["hbm_bytes_per_launch"] for k in parts) if len(parts) == 3 else None
    kms = stage_ms[dom]
    achieved = work / (kms * 1e-3) / (1e9 if unit == "GB/s" else 1e12)
    stage_per_step = {k: v / args.steps for k, v in sm.items()}
    per_kernel = {}
    for k, (kn, bd, wk, pk, un) in kernels.items():
        a = wk / (stage_ms[k] * 1e-3) / (1e9 if un == "GB/s" else 1e12)
        per_kernel[k] = {"kernel": kn, "bound": bd, "ms": round(stage_ms[k], 4), "achieved": round(a, 3),
                         "unit": un, "frac": round(a / pk, 5)}

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
        "dtype": (("f64 (FFT convolution" if conv_used == "fft" else "f32 (convolution") +
                  (", profile values" if args.render_fp64 else "") + ", cell positions, likelihood sums) + f32 (" +
                  ("" if args.render_fp64 else "profile values, ") + "refinement tests, adjoint derivatives)"),
        "data": "synthetic (noiseless model from this GPU path plus seeded SplitMix64/Box-Muller noise)",
        "config": {
            "workload": "400x400 bulge+disk (Sersic n=4 re=3 mag=17 axrat=0.9; Sersic n=1 re=15 mag=16 "
                        "axrat=0.4 ang=60) convolved with a 51x51 Moffat PSF (FWHM 4, con 3); chi^2 over an "
                        "elliptical region (70% of pixels); magzero 30; adjoint gradient of "
                        + ("the 7 metric parameters (xcen, ycen, mag, re, nser, ang, axrat)" if args.grad_slots ==
                           "metric" else "all 11 slots") + " per component",
            "batch": B,
            "drawing_canvas": "450x450",
            "tile_rows": args.tile_rows,
            "profile_values": "fp64" if args.render_fp64 else "fp32",
            "convolution": conv_used,
            "autotune_ms": tune["ms"] if tune else None,
        },
        "adjoint_vs_fd": {
            "adjoint_graph_replay_value": round(graph_value, 3),
            "fd_plain_forward_value": round(fd_value, 3) if fd_value else None,
            "ratio": round(value / fd_value, 2) if fd_value else None,
            "fd_forward_calls_per_lik_grad": 29,
        },
        "stage_ms_per_step": {k: round(v, 4) for k, v in stage_per_step.items()},
        "roofline": {
            "bound": bound,
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": unit,
            "frac": round(achieved / peak, 5),
            "traffic": traffic,
            "kernel": kname,
            "work_per_launch": work,
            "avg_launch_ms": round(kms, 4),
        },
        "roofline_by_kernel": per_kernel,
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
