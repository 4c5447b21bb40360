"""Benchmark of the MI355X deblending path (one GPU).

A step is one ISTA iteration (forward FK operator, blend + residual, adjoint
FK operator, update + soft threshold) over the whole synthetic volume of the
workload, with the problem resident in HBM. Default workload: the learned
warm-start volume, 64 channels x 1024 sources x 2048 samples, windows
(32, 32, 256), shots every 2 s +- 0.5 s.

Reported (one JSON line):
  value        seconds to reach 20 dB deblending SNR with the learned warm
               start: network inference on the device copy of the data, the
               warm-start model, the loop and the output assembly (inputs and
               tables already in HBM), for the smallest n_its of a sweep whose
               output reaches >= 20 dB; config.plain_s is the same for the
               plain zero-start inversion (its own n_its sweep; the threshold
               schedule depends on n_its, so each n_its is a separate run)
  ms_per_step  mean time of the K timed iterations
  roofline     the FK operator: algorithmic bytes 3M + 2T per iteration
               (M model bytes, T trace-buffer bytes) over the device time per
               iteration of its kernels (the spatial and time transforms, the
               update and the shrink, fused or not), timed with HIP events on
               the library stream; traffic is the HBM bytes per iteration of
               the same kernels from the FETCH_SIZE / WRITE_SIZE counter runs
               summarised in profiles/, when that summary matches the workload
  cpu_baseline the oracle (float32, OpenMP over windows, the reference's
               algorithm with a project FFT) timed on this host: the whole
               volume on all affinity cores for n_its 1 and 2 (the measured
               per-iteration cost, extended to the same n_its), and one
               thread on a sub-volume scaled by window count
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FK_KERNELS = ["fwd_spatial", "fwd_time", "adj_time", "adj_spatial", "shrink",
              "fwd_spatial_x3", "adj_spatial_x3"]
OTHER_KERNELS = ["blend", "wavelet_fwd", "wavelet_adj", "norm2"]
PROFILES = os.path.join(ROOT, "profiles")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def geometry_kwargs(g):
    keys = ["volumes", "n_dims", "window_shapes", "coords", "shapes", "shottimes",
            "channels", "trace_types"]
    kw = {k: g[k] for k in keys}
    for k in ("wavelet_lengths", "wavelet_idxs", "wavelets"):
        if k in g:
            kw[k] = g[k]
    return kw


def make_workload(cfg):
    from agdeblend_amd import synth
    if cfg in (1, 2):
        # shots every 4 s +- 1 s: two-fold blending, as in the 2-D workload
        # (C2 is C1's volume with windows (32, 32, 512))
        return synth.WORKLOADS[cfg](interval_s=4.0, dither_s=1.0)
    return synth.WORKLOADS[cfg]()


def traffic_from_summary(workload):
    """HBM bytes per iteration of the FK kernels from the committed counter
    summaries profiles/fk_operator_*_hbm_traffic.md (a line
    'traffic_bytes_per_iteration <workload, spaces as _> <bytes>'), or None."""
    key = workload.replace(" ", "_")
    try:
        names = sorted(os.listdir(PROFILES))
    except OSError:
        return None
    for name in names:
        if not (name.startswith("fk_operator_") and name.endswith("_hbm_traffic.md")):
            continue
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                parts = line.split()
                if (len(parts) == 3 and parts[0] == "traffic_bytes_per_iteration"
                        and parts[1] == key):
                    return float(parts[2])
    return None


def calibration_copy(mib):
    """A plain device copy of mib MiB (read + write), the known-size kernel the
    FETCH_SIZE / WRITE_SIZE counter runs are scaled by."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("calibration copy: 3 x %d MiB" % mib)
    del a, b


def time_to_snr(pm, prob, target_db, sweep, synth):
    """Smallest n_its of the sweep reaching target_db: (seconds, n_its, snr)."""
    results = []
    for n in sweep:
        t0 = time.perf_counter()
        pm.run(1.0, n)
        pm.finish()
        el = time.perf_counter() - t0
        out = pm.output()
        snr = synth.snr(prob, out)
        results.append((n, el, snr))
        log("  n_its %5d  %.3f s  SNR %.2f dB" % (n, el, snr))
        if snr >= target_db:
            return el, n, snr, results
    return None, None, None, results


SWEEPS = {1: "50,100,150,200,300,400,600,800,1200",
          4: "250,300,325,350,375,400,450,500,600,800"}


def warm_start_leg(pm, prob, args, synth):
    """Learned warm start on the workload: per n_its of the sweep, the data
    set again (the output assembly overwrites the device copy), then timed:
    network estimate + agdx_run_warm_dev + output assembly. Also the
    network's own SNR and its inference rate against the MFMA peak."""
    import torch
    from agdeblend_amd import lib as L
    from agdeblend_amd import warmstart as W
    net = W.UNet().load(W.WEIGHTS)
    pm.set_data(prob["blended"])
    est = W.estimate(pm, net)  # first call allocates the activations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    W.estimate(pm, net, est)
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    shape = prob["truth"][0].shape
    snr_net = synth.snr(prob, [est.cpu().numpy().reshape(shape)])
    nvox = est.numel()
    tflops = W.flops_per_voxel() * nvox / t_inf / 1e12
    tflops_x = W.executed_flops_per_voxel() * nvox / t_inf / 1e12
    log("network: inference %.4f s (%.1f TFLOP/s executed, %.1f as the architecture counts), "
        "estimate SNR %.2f dB" % (t_inf, tflops_x, tflops, snr_net))
    res, best = [], None
    for n in [int(v) for v in args.warm_sweep.split(",") if v]:
        pm.set_data(prob["blended"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        W.deblend_warm(pm, net, args.warm_factor, n, est)
        pm.finish()
        el = time.perf_counter() - t0
        snr = synth.snr(prob, pm.output())
        ratio, fell = L.last_warm_ratio(), L.last_warm_fallback()
        # the zero start at the same iteration count, for comparison
        pm.set_data(prob["blended"])
        pm.run(1.0, n)
        snr_plain = synth.snr(prob, pm.output())
        res.append(dict(n_its=n, s=round(el, 4), snr_db=round(snr, 2),
                        zero_start_snr_db=round(snr_plain, 2),
                        guard_ratio=None if ratio is None else round(ratio, 4),
                        fell_back=bool(fell)))
        log("  warm start n_its %3d  %.4f s  SNR %.2f dB (zero start %.2f dB; guard ratio %s%s)"
            % (n, el, snr, snr_plain, "-" if ratio is None else "%.3f" % ratio,
               ", fell back" if fell else ""))
        if snr >= args.target_db:
            best = (el, n, snr)
            break
    return dict(initial_factor=args.warm_factor, weights="agdeblend_amd/weights/unet_c4.npz",
                s_to_target=None if best is None else round(best[0], 4),
                n_its_to_target=None if best is None else best[1],
                snr_db=None if best is None else round(best[2], 2), sweep=res,
                network_snr_db=round(snr_net, 2), inference_s=round(t_inf, 4),
                inference_tflops_executed=round(tflops_x, 1),
                inference_tflops_architecture=round(tflops, 1),
                inference_mfma_frac=round(tflops_x / 2500.0, 4))


def subvolume(prob, axis, n_first):
    """The first n_first entries of space axis `axis` of every patch (same
    seed, same geometry otherwise)."""
    g = dict(prob["geometry"])
    idx = (slice(None),) * axis + (slice(0, n_first),)
    cut = lambda a: [np.asarray(x)[idx] for x in a]  # noqa: E731
    for k in ("shottimes", "channels", "trace_types", "wavelet_idxs"):
        if k in g:
            g[k] = cut(g[k])
    g["shapes"] = [list(s[:axis]) + [n_first] + list(s[axis + 1:]) for s in g["shapes"]]
    sub = dict(prob)
    sub["geometry"] = g
    sub["blended"] = cut(prob["blended"])
    sub["truth"] = cut(prob["truth"])
    return sub


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def fft_backend():
    """The oracle's transforms: its own mixed-radix Stockham and Bluestein
    FFTs (oracle/oracle_fft.c). FFTW, which the reference links, is named
    when this host has it, but the oracle does not use it."""
    import ctypes.util
    have = ctypes.util.find_library("fftw3f") is not None
    return ("project FFT (oracle_fft.c: Stockham radix <= 31, Bluestein beyond); "
            "libfftw3f %s on this host" % ("present, not used" if have else "absent"))


def _oracle_run(oracle, sub, kw, threads):
    oracle.lib(False).set_threads(threads)
    times = []
    for n in (1, 2):
        t0 = time.perf_counter()
        oracle.deblend(data=sub["blended"], initial_factor=1.0, n_its=n, double=False, **kw)
        times.append(time.perf_counter() - t0)
    return times


def cpu_baseline(prob, n_its_target, n_windows_full):
    """The oracle (float32, the reference's algorithm, OpenMP over windows)
    on this host. All affinity cores: the whole volume with n_its = 1 and 2,
    so the per-iteration cost is measured, not scaled. One thread: a
    sub-volume (the first lines of the leading axis) with n_its = 1 and 2,
    scaled by its window count to the whole volume (the work per window is
    the same). value: the all-core time for n_its_target iterations."""
    from oracle import oracle
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    kw = geometry_kwargs(prob["geometry"])
    t1, t2 = _oracle_run(oracle, prob, kw, cores)
    per_it = max(t2 - t1, 1e-9)
    fixed = max(t1 - per_it, 0.0)
    value = fixed + per_it * n_its_target
    # one thread on a bounded sub-volume: the last space axis cut to about
    # 1/16 of its length, never below two windows
    shape = [int(v) for v in prob["geometry"]["shapes"][0]]
    axis = len(shape) - 2
    lead = shape[axis]
    win = int(prob["geometry"]["window_shapes"][0][axis])
    n_sub = min(lead, max(2 * win, lead // 16))
    sub = subvolume(prob, axis, n_sub)
    skw = geometry_kwargs(sub["geometry"])
    m = oracle.Model(double=False, **skw)
    n_windows_sub = m.n_windows
    del m
    s1, s2 = _oracle_run(oracle, sub, skw, 1)
    scale = n_windows_full / max(1, n_windows_sub)
    per_it_1 = max(s2 - s1, 1e-9) * scale
    oracle.lib(False).set_threads(cores)
    return dict(value=round(value, 3), unit="s", cores=cores, kind="port",
                sample="oracle float32 on the whole volume (%d windows), %d threads: n_its=1 "
                       "%.2f s, n_its=2 %.2f s -> %.3f s per iteration, %.2f s fixed; %d "
                       "iterations. One thread on the first %d of %d entries of space axis %d (%d windows): "
                       "n_its=1 %.2f s, n_its=2 %.2f s, x%.2f by windows -> %.2f s per "
                       "iteration" % (n_windows_full, cores, t1, t2, per_it, fixed,
                                      n_its_target, n_sub, lead, axis, n_windows_sub, s1, s2,
                                      scale,
                                      per_it_1),
                per_iteration_s=round(per_it, 4), threads=cores,
                per_iteration_s_1thread=round(per_it_1, 4),
                per_iteration_1thread_kind="scaled from a sub-volume by window count",
                cpu_model=cpu_model(), fft_backend=fft_backend())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--target-db", type=float, default=20.0)
    ap.add_argument("--sweep", type=str, default=None,
                    help="n_its of the plain inversion, ascending")
    ap.add_argument("--warm-sweep", type=str, default="2,3,5,8,12,20,30",
                    help="n_its of the warm-started runs; 2 and up run at least one "
                         "thresholded iteration after the start model")
    ap.add_argument("--warm-factor", type=float, default=0.03)
    ap.add_argument("--no-warm", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default spatial tile instead of the autotuner's choice")
    ap.add_argument("--precision", type=int, default=0,
                    help="spatial DFTs: 0 FP32 register FFTs, 1 bf16 x 3 matrix-core DFTs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-snr", action="store_true")
    ap.add_argument("--no-e2e", action="store_true")
    ap.add_argument("--calib-copy-mib", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU sharding is a design note, not an implementation)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.config == 4 and not args.no_warm:
        # the network runs on torch's HIP runtime, which must initialise the
        # device before the library's
        import torch
        torch.cuda.is_available()
    from agdeblend_amd import agdeblend as agd
    from agdeblend_amd import lib as L
    from agdeblend_amd import synth

    if args.calib_copy_mib > 0:
        calibration_copy(args.calib_copy_mib)

    prob = make_workload(args.config)
    g = prob["geometry"]
    t0 = time.perf_counter()
    prob["blended"] = synth.blend_input(prob, agd.blend)
    log("workload %s; blend %.2f s; input SNR %.2f dB" %
        (prob["name"], time.perf_counter() - t0, synth.snr(prob, prob["blended"])))
    kw = geometry_kwargs(g)
    t0 = time.perf_counter()
    pm = L.Problem(**kw)
    t_setup = time.perf_counter() - t0
    n_x3 = pm.set_spatial_precision(args.precision) if args.precision else 0
    pm.set_data(prob["blended"])
    tune = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        nk, ms = pm.autotune(reps=3)
        tf = pm.time_factors()
        tune = dict(planes_per_task=nk, spatial_ms=[round(t, 4) for t in ms],
                    time_factors=["%dx%d" % f for f in tf],
                    fwd_time_variants=pm.time_variants(),
                    tuning_s=round(time.perf_counter() - t0, 3))
        log("autotune: time factorizations %s (forward variants %s); planes per spatial task "
            "%s (%s ms), %.2f s" % (tune["time_factors"], tune["fwd_time_variants"], nk,
                                    ", ".join("%.3f" % t for t in ms), tune["tuning_s"]))
    n_windows = pm.n_windows
    by = pm.bytes()
    b_fk = 3 * by["M"] + 2 * by["T"]
    log("setup %.2f s; %d windows; M %.3f GB, T %.3f GB, R %.3f GB; B_FK %.3f GB" %
        (t_setup, pm.n_windows, by["M"] / 1e9, by["T"] / 1e9, by["R"] / 1e9, b_fk / 1e9))

    # state for the timed iterations: a short inversion, then fixed-threshold steps
    pm.run(1.0, 3)
    lamb = 0.05 * L.last_lambda0()
    pm.iterate(args.warmup, lamb)
    t0 = time.perf_counter()
    pm.iterate(args.steps, lamb)  # synchronises at the end
    el = time.perf_counter() - t0
    ms_per_step = el / args.steps * 1e3
    log("timed %d iterations: %.3f ms/iteration" % (args.steps, ms_per_step))

    # per-kernel device times (HIP events on the library stream)
    L.timing(True)
    pm.iterate(args.steps, lamb)
    L.timing(False)
    per_launch, per_it = {}, {}
    for k in FK_KERNELS + OTHER_KERNELS:
        ms, n = L.timing_get(k)
        if n:
            per_launch[k] = ms / n
            per_it[k] = ms / args.steps
            log("  %-12s %9.4f ms/launch  %9.4f ms/iteration" % (k, ms / n, ms / args.steps))
    fk_ms = sum(per_it.get(k, 0.0) for k in FK_KERNELS)
    achieved = b_fk / (fk_ms * 1e-3) / 1e9
    traffic = traffic_from_summary(prob["name"] + ("_bf16x3" if args.precision else ""))
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    kernel="FK operator per iteration (%s)" %
                           "+".join(k for k in FK_KERNELS if k in per_it),
                    algorithmic_bytes=b_fk, ms_per_iteration=round(fk_ms, 4),
                    per_kernel_ms_per_launch={k: round(v, 4) for k, v in per_launch.items()},
                    per_kernel_ms_per_iteration={k: round(v, 4) for k, v in per_it.items()})

    plain_s, n20, snr20, sweep_res = None, None, None, []
    if not args.no_snr:
        sw = args.sweep or SWEEPS.get(args.config, SWEEPS[1])
        sweep = [int(s) for s in sw.split(",") if s]
        plain_s, n20, snr20, sweep_res = time_to_snr(pm, prob, args.target_db, sweep, synth)

    warm = None
    if not args.no_snr and not args.no_warm and args.config == 4:
        warm = warm_start_leg(pm, prob, args, synth)

    e2e = None
    if not args.no_snr and not args.no_e2e and n20:
        # the whole agd_deblend call from host buffers (setup, H2D, loop, D2H)
        wav = {k: g[k] for k in ("wavelet_idxs", "wavelets") if k in g}
        t0 = time.perf_counter()
        out = agd.deblend(volumes=g["volumes"], window_shapes=g["window_shapes"],
                          coords=g["coords"], shottimes=g["shottimes"],
                          channels=g["channels"], trace_types=g["trace_types"],
                          initial_factor=1.0, n_its=n20, print_freq=0,
                          data=prob["blended"], **wav)
        e2e = time.perf_counter() - t0
        log("agd_deblend end to end (n_its %d): %.3f s, SNR %.2f dB" %
            (n20, e2e, synth.snr(prob, out)))

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(prob, n20 if n20 else prob["n_its"], n_windows)
        except Exception as e:  # the baseline must not hide the GPU result
            log("cpu baseline failed: %r" % (e,))
    pm.close()

    value = plain_s
    if warm is not None:
        value = warm.get("s_to_target")
        if value is not None and plain_s is not None:
            warm["speedup_vs_plain"] = round(plain_s / value, 2)
    line = dict(
        metric="s to 20 dB deblend SNR (plain vs learned warm start, 1 GPU); FK-op HBM GB/s",
        value=None if value is None else round(value, 4),
        unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
        ms_per_step=round(ms_per_step, 4), higher_is_better=False, scaling="weak",
        vs_baseline=None, dtype="f32", data="synthetic",
        config=dict(workload=prob["name"],
                    value_is="learned warm start" if warm is not None else "plain inversion",
                    plain_s=None if plain_s is None else round(plain_s, 4),
                    n_its_to_target=n20,
                    snr_db=None if snr20 is None else round(snr20, 2),
                    target_db=args.target_db, setup_s=round(t_setup, 3),
                    end_to_end_s=None if e2e is None else round(e2e, 3),
                    sweep=[(n, round(t, 4), round(s, 2)) for n, t, s in sweep_res],
                    parallelism="1 GPU", warm_start=warm, autotune=tune,
                    spatial_dft="bf16x3 MFMA (%d window classes)" % n_x3 if args.precision
                    else "fp32 register FFT"),
        roofline=roofline, cpu_baseline=cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
