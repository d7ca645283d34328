#!/usr/bin/env python3
"""Benchmark of the model-image hot path on one MI355X.

Workload (one "step"): a batch of B likelihood+gradient evaluations of the
400x400 bulge+disk galaxy (Sersic n=4, re=3 + Sersic n=1, re=15) convolved
with a 51x51 Moffat PSF (FWHM 4, con 3), chi^2 likelihood over an elliptical
region of ~70% of the pixels, gradient w.r.t. the 7 parameters of each
component that the metric names.  The B parameter sets are walkers scattered
around the truth.

Printed JSON line: value = batched-adjoint lik+grad evaluations/s; beside it
the plain per-call GPU forward path with central finite differences (29
forward likelihoods per lik+grad), the parity of the GPU result against the
FP64 oracle, the roofline of the dominant kernel from HIP-event timings on the
engine stream over the timed steps (every kernel of the step in
roofline_by_kernel), likelihood-only and forward throughput on the other
configurations, and the FP64 CPU oracle (libprofit's CPU algorithm) timed on
the host cores, all of them and one, on a bounded sample.

--workload c4 runs the 4096-galaxy heterogeneous batch instead.
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, calc, seconds, threads):
    """FP64 oracle: lik+grad by central differences (29 likelihoods), timed on
    `threads` host cores for about `seconds`; plus its model, logL, chi^2 and
    analytic gradient for the parity report."""
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


def calibrate_copy():
    """Copies 256 MiB device to device three times with the library's plain
    copy kernel (the null convolver of profit_amd_convolve_noalloc: each
    launch reads and writes 256 MiB once): the known traffic that scales the
    FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    import profit_amd as P
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    w, h = 8192, 4096  # 32 Mi doubles
    nbytes = w * h * 8
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), nbytes) == 0 and hip.hipMalloc(C.byref(dst), nbytes) == 0
    try:
        assert hip.hipMemset(src, 0, nbytes) == 0
        for _ in range(3):
            P._check(P.lib().profit_amd_convolve_noalloc(P.CONVOLVERS["null"], src, w, h, src, 1, 1, None, dst, 0,
                                                         None))
        assert hip.hipDeviceSynchronize() == 0
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    print(json.dumps({"copy_kernel": "k_noalloc_copy", "copy_bytes_read_per_launch": nbytes,
                      "copy_bytes_written_per_launch": nbytes}))


def sample_counts(eng1, params1, geometry, comps):
    """Profile evaluations per instance at these parameters: the directly
    evaluated pixels (k_render), the cell values of the sub-pixel trees
    (res^2 per tree node; resolution 8 at the Sersic centre pixels), and the
    number of queued (sub-sampled) pixels."""
    cls, nodes = eng1.diagnostics(params1)
    adj = eng1.adjusted_settings(params1)
    _, _, px, py = geometry
    direct = int((cls == 2).sum())
    cells = 0
    for k, c in enumerate(comps):
        sub = cls[k] == 3
        res = np.full(cls[k].shape, int(adj["resolution"][k]))
        if c["type"] == "sersic" and c["nser"] > 1:
            jj, ii = np.nonzero(sub)
            near = (np.abs(ii + 0.5 - (c["xcen"] + px)) < 1) & (np.abs(jj + 0.5 - (c["ycen"] + py)) < 1)
            res[jj[near], ii[near]] = 8
        cells += int((nodes[k][sub].astype(np.int64) * res[sub] ** 2).sum())
    return direct, cells, int((cls == 3).sum())


def fft_work(dw, dh, w, h, n1, n2):
    """FLOPs each FFT kernel executes per instance (5 N log2 N per complex
    transform of length N; two real rows per complex transform; the spectrum
    product 6 FLOPs per entry), likelihood stage then adjoint stage, and the
    stages' algorithmic HBM bytes (likelihood: canvas 4 B per canvas pixel and
    data, 1/sigma^2, region and w: 13 B per image pixel; adjoint: w in and G
    out, 4 B each)."""
    row = 5.0 * n1 * np.log2(n1)
    col = (n1 // 2 + 1) * (2 * 5.0 * n2 * np.log2(n2) + 6.0 * n2)
    pairs = lambda r: (r + 1) // 2
    flops = {"like_rows_fwd": pairs(dh) * row, "like_cols": col, "like_rows_inv": pairs(h) * row,
             "adj_rows_fwd": pairs(h) * row, "adj_cols": col, "adj_rows_inv": pairs(dh) * row}
    stage_bytes = {"conv_like": 4.0 * dw * dh + 13.0 * w * h, "adjoint_conv": 4.0 * w * h + 4.0 * dw * dh}
    # each kernel's algorithmic HBM bytes: the half-spectrum rows it reads and
    # writes (16 B per entry, N1/2 + 1 entries a row) and its image data
    m = 16.0 * (n1 // 2 + 1)
    kernel_bytes = {"like_rows_fwd": 4.0 * dw * dh + m * dh, "like_cols": m * (dh + h),
                    "like_rows_inv": m * h + 20.0 * w * h, "adj_rows_fwd": 4.0 * w * h + m * h,
                    "adj_cols": m * (h + dh), "adj_rows_inv": m * dh + 4.0 * dw * dh}
    return flops, stage_bytes, kernel_bytes


def time_calls(fn, reps):
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    return reps / (time.perf_counter() - t0)


def other_configs(P, W, reps):
    """Likelihood-only (forward + likelihood) and lik+grad throughput on
    configurations 1 and 5, and the per-call model image of Model::evaluate
    (profit_amd_make_model, host arrays in and out, as R's profitMakeModel
    calls it) on configurations 1 and 2."""
    out = {}

    def model_fn(w, h, psf, comps, magzero):
        return P.make_model(w, h, gpu_params(P, [comps]), psf=psf, magzero=magzero)["image"]

    # config 1: 100x100 Sersic, 25x25 Gaussian PSF, brute force, chi^2; a batch of 1024
    wl1 = W.config1(model_fn, magzero=30.0)
    eng = P.Engine()
    g = eng.add_galaxy(100, 100, wl1["data"], wl1["sigma"], psf=wl1["psf"], region=wl1["region"], magzero=30.0)
    eng.set_conv_method("brute")
    eng.set_grad_slots(("xcen", "ycen", "mag", "ang", "axrat", "s0", "s1"))
    B1 = 1024
    eng.set_batch([g] * B1, [1] * B1)
    p1 = gpu_params(P, walkers(wl1["components"], B1, seed=21))
    like1 = time_calls(lambda: eng.like(p1), reps) * B1
    grad1 = time_calls(lambda: eng.like_grad(p1), reps) * B1
    eng.close()
    out["config1"] = {"workload": "100x100 Sersic n=2.5 re=8 convolved with a 25x25 Gaussian (FWHM 3), brute "
                                  "force, chi^2, magzero 30, batch 1024",
                      "like_evals_per_s": round(like1, 1), "lik_grad_evals_per_s": round(grad1, 1)}
    prof1 = gpu_params(P, [wl1["components"]])
    mm1 = time_calls(lambda: P.make_model(100, 100, prof1, psf=wl1["psf"], magzero=30.0), reps)
    out["config1"]["make_model_calls_per_s"] = round(mm1, 1)

    # config 2 per call through the model path (the batched form is the main metric)
    wl2 = W.config2()
    prof2 = gpu_params(P, [wl2["components"]])
    for conv in ("brute", "fft"):
        mm2 = time_calls(lambda: P.make_model(400, 400, prof2, psf=wl2["psf"], magzero=30.0, convolver=conv), reps)
        out.setdefault("config2", {})[f"make_model_calls_per_s_{conv}"] = round(mm2, 1)

    # config 5: 2048x2048 field of 500 Sersics, 101x101 Moffat PSF, one instance
    wl5 = W.config5(model_fn)
    eng = P.Engine()
    g = eng.add_galaxy(2048, 2048, wl5["data"], wl5["sigma"], psf=wl5["psf"], region=wl5["region"], magzero=30.0)
    eng.set_grad_slots(("xcen", "ycen", "mag", "ang", "axrat", "s0", "s1"))
    eng.set_batch([g], [len(wl5["components"])])
    p5 = gpu_params(P, [wl5["components"]])
    tune = eng.autotune(p5, reps=1)
    like5 = time_calls(lambda: eng.like(p5), reps)
    grad5 = time_calls(lambda: eng.like_grad(p5), reps)
    eng.close()
    out["config5"] = {"workload": "2048x2048, 500 Sersics (max_recursions 4), 101x101 Moffat PSF, chi^2, one "
                                  "instance", "convolution": tune["kinds"][0],
                      "like_evals_per_s": round(like5, 2), "lik_grad_evals_per_s": round(grad5, 2)}
    return out


def config4_leg(P, W, steps, warmup):
    """The 4096-galaxy heterogeneous batch (config 4): lik+grad of every
    instance per step, with the autotuner's per-galaxy convolution kernels."""
    t0 = time.perf_counter()
    gals = W.config4_galaxies(4096, seed=4)
    for g in gals:
        g["calc"] = P.adjust_mask(g["region"], (g["width"], g["height"]), g["psf"])
        g["data"] = np.zeros((g["height"], g["width"]))
        g["sigma"] = np.ones((g["height"], g["width"]))

    def add(eng, g):
        return eng.add_galaxy(g["width"], g["height"], g["data"], g["sigma"], psf=g["psf"], calcregion=g["calc"],
                              region=g["region"], magzero=30.0)

    truth = P.profile_array([p for g in gals for p in
                             [P.make_profile(c["type"], convolve=True, **{k: v for k, v in c.items() if k != "type"})
                              for c in g["components"]]])
    meng = P.Engine()
    ids = [add(meng, g) for g in gals]
    meng.set_batch(ids, [len(g["components"]) for g in gals])
    models = meng.model_images(truth)
    meng.close()
    W.config4_attach_data(gals, [m.astype(np.float64) for m in models])
    eng = P.Engine()
    ids = [add(eng, g) for g in gals]
    eng.set_batch(ids, [len(g["components"]) for g in gals])
    tune = eng.autotune(truth, reps=1)
    setup_s = time.perf_counter() - t0
    for _ in range(warmup):
        eng.like_grad(truth)
    t1 = time.perf_counter()
    for _ in range(steps):
        eng.like_grad(truth)
    t2 = time.perf_counter()
    # stage times from separate steps with per-stage events on
    eng.set_timing(True)
    eng.reset_stats()
    for _ in range(steps):
        eng.like_grad(truth)
    stats = eng.stats()
    eng.set_timing(False)
    eng.close()
    n = len(gals)
    kinds = tune["kinds"]
    return {
        "workload": "4096 galaxies, sides U(32, 256) px, 1-3 components (Sersic, Moffat, Ferrer, King, core-Sersic), "
                    "own Gaussian or Moffat PSF 15-41 px, chi^2 over 0.48-axis ellipses, magzero 30",
        "lik_grad_evals_per_s": round(n * steps / (t2 - t1), 1),
        "ms_per_step": round((t2 - t1) * 1e3 / steps, 3),
        "stage_ms_per_step": {k: round(v / steps, 4) for k, v in stats["stage_ms"].items()},
        "convolution_kinds": {k: kinds.count(k) for k in sorted(set(kinds))},
        "autotune_ms": tune["ms"],
        "setup_s": round(setup_s, 1),
        "components": int(sum(len(g["components"]) for g in gals)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--workload", choices=("c2", "c4"), default="c2",
                    help="c2: the metric's 400x400 bulge+disk batch; c4: the 4096-galaxy heterogeneous batch")
    ap.add_argument("--fd-evals", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the likelihood-only / per-call figures of configurations 1, 2 and 5")
    ap.add_argument("--tile-rows", type=int, default=8)
    ap.add_argument("--render-fp64", type=int, nargs="?", const=1, default=0, choices=(0, 1, 2),
                    help="model precision: 1 (the flag alone) profile values (direct pixels, recursion leaves) in "
                         "FP64 instead of FP32; 2 also an FP64 canvas and the FP64 FFT for every PSF")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 256 MiB device-to-device three times with the library's copy kernel (the "
                         "reference kernel for scaling the FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--conv", choices=("auto", "brute", "fft", "brute-bf16x3"), default="auto",
                    help="PSF convolution method (auto: the autotuner's choice among brute and fft)")
    ap.add_argument("--grad-slots", choices=("metric", "all"), default="metric",
                    help="gradient computed per component: 'metric' = the 7 parameters of the 14-parameter "
                         "bulge+disk gradient (xcen, ycen, mag, re, nser, ang, axrat); 'all' adds box and "
                         "the unused shape slots")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import profit_amd as P
    from profit_amd import workloads as W

    if args.workload == "c4":
        r = config4_leg(P, W, args.steps, args.warmup)
        print(json.dumps({"metric": "lik+grad evals/s, config 4 heterogeneous batch", "value": r["lik_grad_evals_per_s"],
                          "unit": "lik+grad evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": r["ms_per_step"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": None, "dtype": "f32 render and brute convolution, f64 FFT and sums",
                          "data": "synthetic (GPU model plus seeded noise)", "config": r}))
        return

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
        log(f"autotune: {tune['kinds'][0]} ({tune['ms']})")
    conv_used = tune["kinds"][0] if tune else args.conv
    log(f"batch {B}: warmup {args.warmup}")
    for _ in range(args.warmup):
        eng.like_grad(params)

    # timed region: eager launch sequence with HIP events on the engine stream
    # between the stages and between the FFT kernels
    eng.set_timing(True)
    eng.reset_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ll, grad = eng.like_grad(params)
    t1 = time.perf_counter()
    stats = eng.stats()
    fft_kernel_ms = eng.fft_stats()
    eng.set_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    log(f"adjoint: {value:.1f} lik+grad/s, {ms_per_step:.2f} ms/step")
    log("stage ms per step: " + ", ".join(f"{k} {v / args.steps:.3f}" for k, v in stats["stage_ms"].items()))
    log("FFT kernels ms per step: " + ", ".join(f"{k} {v:.3f}" for k, v in fft_kernel_ms.items()))

    # likelihood only (no gradient) on the same batch
    like_value = time_calls(lambda: eng.like(params), args.steps) * B
    # graph replay of the same batch (production path)
    graph_value = time_calls(lambda: eng.like_grad(params), args.steps) * B

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
        log(f"plain forward + FD: {fd_value:.1f} lik+grad/s")

    # per-kernel rooflines (one launch = the whole batch); algorithmic work per
    # unit of each kernel family:
    #   rendering (k_render, the recursion walkers, the adjoint render): 5 B per
    #     canvas pixel (4 B FP32 write + 1 B mask read), reported against HBM;
    #     their limit is the VALU / transcendental rate, so profile
    #     evaluations per second are given beside it
    #   FFT kernels: the FP64 FLOPs each executes; the stage's algorithmic bytes
    #   brute-force convolution: 2 K^2 FLOP per output pixel (FP32 vector peak)
    sm = stats["stage_ms"]
    ncalls = max(1, stats["stage_calls"]["subsample"])
    stage_ms = {k: v / ncalls for k, v in sm.items()}
    geom = eng.geometry(g)
    dw, dh, _, _ = geom
    K = wl["psf"].shape[0]
    n1 = max(8, 1 << int(np.ceil(np.log2(dw + K))))
    n2 = max(8, 1 << int(np.ceil(np.log2(dh + K))))
    render_bytes = B * dw * dh * 5.0
    direct, cells, queued = sample_counts(eng1, gpu_params(P, [wl["components"]]), geom, wl["components"])
    traffic = {}
    tfile = os.path.join(ROOT, "profiles", f"traffic_config2_batch{B}.json")
    if os.path.exists(tfile) and conv_used == "fft" and not args.render_fp64:
        with open(tfile) as f:
            traffic = {k: v["hbm_bytes_per_launch"] for k, v in json.load(f)["kernels"].items()}
    kernels = {}

    def add_kernel(key, name, ms, bound, work, peak, unit, extra=None):
        a = work / (ms * 1e-3) / (1e9 if unit == "GB/s" else 1e12) if ms > 0 else 0.0
        e = {"kernel": name, "bound": bound, "ms": round(ms, 4), "achieved": round(a, 3), "peak": peak, "unit": unit,
             "frac": round(a / peak, 5), "work_per_launch": work}
        t = traffic.get(name)
        if t:
            e["traffic"] = t
            e["traffic_GBs"] = round(t / (ms * 1e-3) / 1e9, 1)
            e["traffic_frac"] = round(t / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
        if extra:
            e.update(extra)
        kernels[key] = e

    add_kernel("render", "k_render<0, false>", stage_ms["render"], "hbm", render_bytes, PEAK_HBM_GBS, "GB/s",
               {"profile_evals_per_s": round(B * direct / (stage_ms["render"] * 1e-3), 0),
                "direct_pixels_per_instance": direct})
    # the forward walk of a gradient evaluation: cell values, refinement tests
    # and each cell's derivatives into the pixel's Jacobian
    add_kernel("subsample", "k_subsample_root<0, true> + k_subsample<0, true>", stage_ms["subsample"], "hbm",
               render_bytes, PEAK_HBM_GBS,
               "GB/s", {"cell_value_and_derivative_evals_per_s": round(B * cells / (stage_ms["subsample"] * 1e-3), 0),
                        "cells_per_instance": cells})
    add_kernel("adjoint_render", "k_adj_render<0, *>", stage_ms["adjoint_render"], "hbm", render_bytes, PEAK_HBM_GBS,
               "GB/s", {"derivative_evals_per_s": round(B * direct / (stage_ms["adjoint_render"] * 1e-3), 0)})
    # the adjoint of the sub-sampled pixels: a weighted reduction of the
    # Jacobians, 48 B per queued entry read, against HBM
    add_kernel("adjoint_subsample", "k_adj_queue<0>", stage_ms["adjoint_subsample"], "hbm", B * queued * 60.0,
               PEAK_HBM_GBS, "GB/s", {"queued_entries_per_instance": queued,
                                      "note": "work: 60 B per queued pixel (48 B Jacobian and value, 8 B entry, 4 B G)"})
    if conv_used == "fft":
        flops, stage_bytes, kbytes = fft_work(dw, dh, 400, 400, n1, n2)
        names = {"like_rows_fwd": "k_fft_rows_fwd", "like_cols": "k_fft_cols", "like_rows_inv": "k_fft_rows_inv<3>",
                 "adj_rows_fwd": "k_fft_rows_fwd", "adj_cols": "k_fft_cols", "adj_rows_inv": "k_fft_rows_inv<2>"}
        for key, kname in names.items():
            ms = fft_kernel_ms[key]
            # against HBM: the half-spectrum rows each kernel streams (and its
            # image data); its executed FP64 FLOPs beside it
            add_kernel(key, kname, ms, "hbm", B * kbytes[key], PEAK_HBM_GBS, "GB/s",
                       {"fp64_tflops": round(B * flops[key] / (ms * 1e-3) / 1e12, 3),
                        "fp64_frac": round(B * flops[key] / (ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 5)})
            kernels[key]["note"] = ("work: the half-spectrum rows read and written (16 B an entry) and the image "
                                    "data; fp64_tflops: executed FP64 FLOPs (5 N log2 N per transform)")
        for stage in ("conv_like", "adjoint_conv"):
            add_kernel(stage, "FFT stage (3 kernels)", stage_ms[stage], "hbm", B * stage_bytes[stage], PEAK_HBM_GBS,
                       "GB/s")
    else:
        peak = PEAK_BF16_TFLOPS if conv_used == "brute-bf16x3" else PEAK_FP32_TFLOPS
        kname = "k_conv_mfma" if conv_used == "brute-bf16x3" else "k_conv"
        bound = "mfma" if conv_used == "brute-bf16x3" else "hbm"
        add_kernel("conv_like", f"{kname}<CONV_LIKE>", stage_ms["conv_like"], bound, stats["conv_flops"], peak,
                   "TFLOP/s")
        add_kernel("adjoint_conv", f"{kname}<CONV_ADJ>", stage_ms["adjoint_conv"], bound, stats["adjoint_flops"], peak,
                   "TFLOP/s")
        for k in ("conv_like", "adjoint_conv"):
            kernels[k]["note"] = "work: 2 K^2 FLOP per output pixel; FP32 VALU" if bound == "hbm" else \
                "work: 2 K^2 FLOP per output pixel (three split-BF16 products issued per MAC)"
    # the dominant single kernel of the step (stages of several kernels excluded)
    single = [k for k in kernels if not kernels[k]["kernel"].startswith("FFT stage") and "*" not in kernels[k]["kernel"]
              and "+" not in kernels[k]["kernel"]]
    dom = max(single, key=lambda k: kernels[k]["ms"])
    d = kernels[dom]
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
        "dtype": (("f64 (FFT convolution" if conv_used == "fft" else "f32 (convolution") +
                  (", profile values" if args.render_fp64 else "") + (", canvas" if args.render_fp64 == 2 else "") +
                  ", cell positions, likelihood sums) + f32 (" +
                  ("" if args.render_fp64 else "profile values, ") + "refinement tests, adjoint derivatives)"),
        "data": "synthetic (noiseless model from this GPU path plus seeded SplitMix64/Box-Muller noise)",
        "config": {
            "workload": "400x400 bulge+disk (Sersic n=4 re=3 mag=17 axrat=0.9; Sersic n=1 re=15 mag=16 "
                        "axrat=0.4 ang=60) convolved with a 51x51 Moffat PSF (FWHM 4, con 3); chi^2 over an "
                        "elliptical region (70% of pixels); magzero 30; adjoint gradient of "
                        + ("the 7 metric parameters (xcen, ycen, mag, re, nser, ang, axrat)" if args.grad_slots ==
                           "metric" else "all 11 slots") + " per component",
            "batch": B,
            "drawing_canvas": f"{dw}x{dh}",
            "tile_rows": args.tile_rows,
            "profile_values": "fp64" if args.render_fp64 else "fp32",
            "canvas": "fp64" if args.render_fp64 == 2 else "fp32",
            "convolution": conv_used,
            "autotune_ms": tune["ms"] if tune else None,
        },
        "adjoint_vs_fd": {
            "adjoint_graph_replay_value": round(graph_value, 3),
            "fd_plain_forward_value": round(fd_value, 3) if fd_value else None,
            "ratio": round(value / fd_value, 2) if fd_value else None,
            "fd_forward_calls_per_lik_grad": 29,
        },
        "like_only_evals_per_s": round(like_value, 1),
        "stage_ms_per_step": {k: round(v, 4) for k, v in stage_per_step.items()},
        "roofline": {
            "bound": d["bound"],
            "achieved": d["achieved"],
            "peak": d["peak"],
            "unit": d["unit"],
            "frac": d["frac"],
            "traffic": d.get("traffic"),
            "kernel": d["kernel"],
            "work_per_launch": d["work_per_launch"],
            "avg_launch_ms": d["ms"],
        },
        "roofline_by_kernel": kernels,
    }
    for k in ("cell_evals_per_s", "cell_derivative_evals_per_s", "profile_evals_per_s"):
        if k in d:
            out["roofline"][k] = d[k]

    # FP64 CPU oracle leg: baseline timing and parity of the GPU result
    if not args.no_cpu_baseline:
        threads = os.cpu_count() or 1
        cb = cpu_baseline(wl, calc, args.cpu_seconds, threads)
        cb1 = cpu_baseline(wl, calc, 1.0, 1)  # one lik+grad (29 likelihoods) on one core
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
        reg = wl["region"].astype(bool)
        chi_gpu = float(np.sum(((wl["data"] - gmodel) / wl["sigma"])[reg] ** 2))
        chi_ref = float(np.sum(((wl["data"] - ref) / wl["sigma"])[reg] ** 2))
        out["parity_vs_fp64_oracle"] = {
            "max_rel_pixel_err": pix_err,
            "pixels_checked": int(sel.sum()),
            "abs_logL_err": float(abs(gll[0] - cb["ll"])),
            "abs_chisq_err": abs(chi_gpu - chi_ref),
            "chisq": chi_ref,
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
            "one_core": {"value": round(cb1["value"], 5), "cores": 1,
                         "sample": f"{cb1['n']} lik+grad evaluation(s) in {cb1['elapsed']:.1f} s"},
            "cpu_model": cpu_model(),
        }
    eng1.close()
    eng.close()
    if not args.no_other_configs:
        out["other_configs"] = other_configs(P, W, reps=max(3, args.steps // 2))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
