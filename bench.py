#!/usr/bin/env python3
"""Benchmark: OPAL-T space-charge PIC step on one MI355X.

Workload (the configuration the headline metric is quoted on): 1e8
macroparticles, 6D Gaussian (sigma 1 mm, normalised momentum spread 1e-3)
with 10% of the positions uniform in the 5-sigma ball, 250 MeV protons,
10 nC, 256^3 grid (512^3 doubled-grid convolution), drift, dT = 1e-11 s.
A step is one ParallelTTracker step over the resident bunch: push, beam-frame
rotation, boundp, CIC scatter, IGF Green's function + Hockney convolution,
E = -Grad(phi), gather, kick, push, reference-frame update, statistics.

value = mixed-precision particle-steps/s (FP32 FFT/convolution/gather, FP64
IGF, moments and particle state).  The plain FP64 step is timed on the same
bunch for the speed ratio, and the E-field / emittance errors of the mixed
step against the FP64 step are reported.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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

HBM_PEAK_GBS = 8000.0
MP_EV = 0.93827204e+00 * 1e9


def gamma_betagamma(ekin_mev):
    g = 1.0 + ekin_mev * 1e6 / MP_EV
    return g, math.sqrt(g * g - 1.0)


CONFIGS = {
    # the metric's configuration
    "c3": dict(N=100_000_000, n=256, ekin=250.0, q=10e-9, halo=0.1, Bz=0.0, seed=3),
    # plain FP64 drop-in configuration, used for smaller runs
    "c2": dict(N=10_000_000, n=128, ekin=1.0, q=10e-9, halo=0.0, Bz=0.5, seed=2),
    "c1": dict(N=100_000, n=32, ekin=1.0, q=1e-9, halo=0.0, Bz=0.0, seed=1),
}

# energy-binned configuration: 5e7 particles in 10 bins of populations in the
# ratio 2:1 from bin to bin, 128^3 grid per bin with its image charge, solenoid
# plus RF cavity field, mixed precision; the bunch centred 6 sigma in front of
# the cathode plane z = 0 that the image charges mirror it in
C5 = dict(N=50_000_000, n=128, ekin=1.0, q=10e-9, nbins=10, sigma_p=5e-3, Bz=1.0, Ez0=1e6, rf=50e6,
          seed=5, z0=6e-3)


def make_ctx(sc, cfg, precision, sort_interval):
    n = cfg["n"]
    c = sc.SpaceChargeContext((n, n, n), precision=precision, max_particles=cfg["N"],
                              sort_interval=sort_interval)
    g, bg = gamma_betagamma(cfg["ekin"])
    c.generate_bunch(cfg["N"], cfg["seed"], (1e-3, 1e-3, 1e-3), (1e-3, 1e-3, 1e-3), bg, cfg["q"],
                     halo_frac=cfg["halo"], halo_radius=5.0)
    c.track_init(1e-11, charge=1.0, mass_eV=MP_EV, Bz=cfg["Bz"], ref_p=(0.0, 0.0, bg))
    return c


def timed_run(c, steps, warmup):
    if warmup:
        c.track(warmup)
    sc_sync = c._c  # noqa: F841
    from src_amd.opalsc import lib
    lib().opalsc_synchronize(c._h)
    t0 = time.perf_counter()
    stats = c.track(steps)
    lib().opalsc_synchronize(c._h)
    t1 = time.perf_counter()
    return (t1 - t0), stats


def geometric_bins_by_pz(pz, p0, sigma, nbins):
    """Energy bins from the longitudinal momentum: thresholds at the Gaussian
    quantiles that give bin k the fraction 2^-k / sum_j 2^-j (lowest pz in bin 0)."""
    from statistics import NormalDist
    w = 0.5 ** np.arange(nbins)
    frac = np.cumsum(w)[:-1] / w.sum()
    edges = np.array([p0 + sigma * NormalDist().inv_cdf(float(f)) for f in frac])
    return np.searchsorted(edges, pz, side="right").astype(np.int32)


def c5_context(sc, precision, sort_interval=20):
    """The energy-binned configuration's bunch in its 10 bins, ready to track."""
    cfg = C5
    N, n = cfg["N"], cfg["n"]
    c = sc.SpaceChargeContext((n, n, n), precision=precision, max_particles=N, sort_interval=sort_interval)
    g, bg = gamma_betagamma(cfg["ekin"])
    sp = cfg["sigma_p"]
    c.generate_bunch(N, cfg["seed"], (1e-3, 1e-3, 1e-3), (sp, sp, sp), bg, cfg["q"])
    R, P, Q = c.get_bunch()
    R[:, 2] += cfg["z0"]
    c.set_bunch(R, P, Q)
    del R, Q
    bins = geometric_bins_by_pz(P[:, 2], bg, sp, cfg["nbins"])
    del P
    c.set_bins(bins, cfg["nbins"], cfg["q"] / N, image=True)
    del bins
    c.track_init(1e-11, charge=1.0, mass_eV=MP_EV, Bz=cfg["Bz"], Ez0=cfg["Ez0"], rf_freq=cfg["rf"],
                 ref_p=(0.0, 0.0, bg))
    return c


def run_c5(sc, steps, warmup, sort_interval, log):
    """The energy-binned step (one Poisson solve per bin and one for its image)."""
    cfg = C5
    N, n = cfg["N"], cfg["n"]
    sp = cfg["sigma_p"]
    c = c5_context(sc, sc.MIXED, sort_interval)
    counts, _ = c.get_bins()
    t, st = timed_run(c, steps, warmup)
    ms = 1e3 * t / steps
    c.profile_enable(True)
    c.track(max(sort_interval, 3))
    prof = c.profile_read()
    c.profile_enable(False)
    _, gammas = c.get_bins()
    c.close()
    log("[bench] c5 binned: %.3f ms/step" % ms)
    return {"workload": "5e7 macroparticles in %d energy bins (populations %s), %d^3 grid per bin "
                        "(%d^3 doubled FFT) plus its image charge, %g MeV protons, sigma_p %g, Bz %g T, "
                        "Ez %g V/m at %g MHz, bunch centred %g mm in front of the image plane, mixed precision, "
                        "sort every %d steps" %
                        (cfg["nbins"], "/".join(str(int(x)) for x in counts), n, 2 * n, cfg["ekin"], sp,
                         cfg["Bz"], cfg["Ez0"], cfg["rf"] / 1e6, 1e3 * cfg["z0"], sort_interval),
            "ms_per_step": ms, "particle_steps_per_s": N * steps / t, "steps": steps,
            "poisson_solves_per_step": 2 * int((counts > 0).sum()),
            "roofline_step": step_roofline(step_bytes(N, n, 4, solves=2 * int((counts > 0).sum())), ms),
            "bin_gammas": [float(x) for x in gammas],
            "emittance_last": [float(x) for x in st[-1]["eps_norm"]],
            "kernels": {k: {"avg_ms": ms_ / nl, "launches": nl} for k, (ms_, nl) in prof.items() if nl}}


def c4_specs():
    """The 64 independent bunches: (N, n, sigma, beta*gamma) with particle
    counts log-uniform in [1e4, 1e6], grids from {16, 32, 48, 64}^3, sigma
    uniform in [0.5, 2] mm and 1-100 MeV protons, from one seeded generator."""
    rng = np.random.default_rng(4)
    specs = []
    for k in range(64):
        N = int(10 ** rng.uniform(4, 6))
        n = int(rng.choice([16, 32, 48, 64]))
        sig = float(rng.uniform(0.5e-3, 2e-3))
        _, bg = gamma_betagamma(float(rng.uniform(1.0, 100.0)))
        specs.append((N, n, sig, bg))
    return specs


def c4_context(sc, k, spec, precision, sort_interval=20):
    """Bunch k of the 64-bunch set (seed 1000 + k, 1 nC) ready to track,
    re-sorted by cell every sort_interval steps."""
    N, n, sig, bg = spec
    c = sc.SpaceChargeContext((n, n, n), precision=precision, max_particles=N, sort_interval=sort_interval)
    c.generate_bunch(N, 1000 + k, (sig, sig, sig), (1e-3, 1e-3, 1e-3), bg, 1e-9)
    c.track_init(1e-11, charge=1.0, mass_eV=MP_EV, ref_p=(0.0, 0.0, bg))
    return c


def run_c4(sc, steps, warmup, log, only_fused=False):
    """64 independent bunches (c4_specs, per-bunch seeds, mixed precision),
    stepped one bunch after another with opalsc_track, then together with
    opalsc_step_batched without and with the HIP graph."""
    specs = c4_specs()
    total = sum(s[0] for s in specs)

    def make():
        return [c4_context(sc, k, s, sc.MIXED) for k, s in enumerate(specs)]

    from src_amd.opalsc import lib
    out = {"workload": "64 bunches, %d particles in total (1e4-1e6 each), grids 16/32/48/64^3, mixed precision"
                       % total, "steps": steps}
    t_seq = float("nan")
    if not only_fused:
        cs = make()
        for c in cs:
            c.track(warmup, want_stats=False)
        t0 = time.perf_counter()
        for c in cs:
            c.track(steps, want_stats=False)
        t_seq = time.perf_counter() - t0
        for c in cs:
            c.close()
    # fused: one launch per pass for all bunches of a kernel class / grid shape
    # (re-sorts every 20 steps launched directly); per_bunch: one graph branch
    # per bunch (no re-sort)
    modes = (("per_bunch_graph", False, True), ("batched_streams", True, False), ("batched_graph", True, True))
    if only_fused:
        modes = modes[2:]
    for mode, fused, graph in modes:
        cs = make()
        with sc.Batch(cs, fused=fused) as b:
            b.step(warmup, graph=graph, want_stats=False)
            t0 = time.perf_counter()
            b.step(steps, graph=graph, want_stats=False)
            t = time.perf_counter() - t0
        for c in cs:
            c.close()
        out[mode + "_ms_per_step"] = 1e3 * t / steps
        out[mode + "_particle_steps_per_s"] = total * steps / t
    c4_bytes = sum(step_bytes(Nb, nb, 4) for Nb, nb, _, _ in specs)
    for mode, _, _ in modes:
        out[mode + "_roofline_step"] = step_roofline(c4_bytes, out[mode + "_ms_per_step"])
    out["sequential_ms_per_step"] = 1e3 * t_seq / steps
    out["sequential_particle_steps_per_s"] = total * steps / t_seq
    log("[bench] c4: " + ", ".join("%s %.2f ms/step" % (k[:-12], v) for k, v in out.items()
                                   if k.endswith("_ms_per_step")))
    return out


def step_bytes(N, n, wbytes, solves=1):
    """Algorithmic HBM bytes of one step of the reference algorithm: per
    particle R and P read and written (96 B; the bunch's equal charges are one
    number), and per Poisson solve on an n^3 grid: the int64 charge written and
    read, phi and the 3-component E written and read, the doubled-grid
    transforms of rho and of G (M = (2n)^3 real points, Mc = (2n)^2 (n+1)
    complex), the real G, the spectral product, the inverse transform and the
    (n+1)^3 corner table; wbytes is the word size of the grids (8 FP64, 4
    mixed)."""
    n3, M, Mc = n ** 3, (2 * n) ** 3, (2 * n) ** 2 * (n + 1)
    grid = (n3 * (2 * 8 + 2 * wbytes + 6 * wbytes) + 2 * Mc * 2 * wbytes + 2 * M * wbytes + 3 * Mc * 2 * wbytes
            + Mc * 2 * wbytes + (n + 1) ** 3 * 8)
    return N * 96 + solves * grid


def step_roofline(bytes_per_step, ms):
    a = bytes_per_step / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
            "bytes_per_step": bytes_per_step}


def profile_kernels(c, steps, ab):
    """Per-kernel average durations (HIP events on the library's stream) over
    `steps` tracker steps, with GB/s and fraction of peak where ab gives the
    kernel's algorithmic bytes per launch."""
    c.profile_enable(True)
    c.track(steps)
    prof = c.profile_read()
    c.profile_enable(False)
    kernels = {}
    for name, (ms, nl) in prof.items():
        if nl == 0:
            continue
        avg = ms / nl
        k = {"avg_ms": avg, "launches": nl}
        if name in ab:
            k["algorithmic_GBs"] = ab[name] / (avg * 1e-3) / 1e9
            k["frac_of_peak"] = k["algorithmic_GBs"] / HBM_PEAK_GBS
        kernels[name] = k
    return kernels


def algorithmic_bytes(cfg, wbytes):
    """Per-launch algorithmic HBM bytes of each kernel class (one step)."""
    N, n = cfg["N"], cfg["n"]
    n3, K0 = n ** 3, n + 1
    cw = 2 * wbytes  # complex word
    return {
        # first step of a call: read x,y,z,px,py,pz; write x,y,z
        "push_rotate_bounds": N * 72,
        # later steps of a call take their bounds from the previous gather pass's
        # chunk statistics and rotate only the chunks that can hold an extreme:
        # no fixed per-particle traffic, reported by time only
        # read x,y,z (the synthetic bunch's particles all carry the same charge,
        # kept as one number); read+write int64 rho once
        "deposit": N * 24 + n3 * 16,
        # read+write x,y,z,px,py,pz; read the E grid (4 words per cell) once
        "gather_kick_push": N * 96 + n3 * 4 * wbytes,
        # the same applying a re-sort: also the permutation and the ids (4 B each, read and written)
        "gather_kick_push_resort": N * 108 + n3 * 4 * wbytes,
        # x pass: int64 rho -> A, and A -> phi (two launches; per-launch average)
        "fft_x": (n3 * 8 + K0 * n * n * cw + K0 * n * n * cw + n3 * wbytes) / 2,
        # y passes: A -> B, B -> A (two launches)
        "fft_y": (K0 * n * n * cw + K0 * 2 * n * n * cw) * 2 / 2,
        # z conv: B in/out + Ghat
        "fft_z_conv": 2 * K0 * 2 * n * n * cw + (n + 1) ** 3 * wbytes,
        "grad": n3 * wbytes + n3 * 4 * wbytes,
        "zero_rho": n3 * 8,
    }


# profiler kernel class -> kernel symbol of the mixed-precision step
KERNEL_SYMBOL = {
    "gather_kick_push": "k_track_gamma_mixed_os<true,0,false,false,false>",
    "deposit": "k_deposit_lds<3,false,true,4864,false>",
    "rotate_bounds": "k_bounds_scan<false>",
    "push_rotate_bounds": "k_track_alpha<false,true,false>",
    "fft_z_conv": "k_strided<float,8,2>",
    "grad": "k_grad<float>",
}


def measured_traffic(config, kernel):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated) of
    `kernel`, from the committed rocprofv3 summary of this workload; None if absent."""
    path = os.path.join(ROOT, "profiles", "%s_mixed_hbm_traffic.txt" % config)
    sym = KERNEL_SYMBOL.get(kernel)
    if sym is None or not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 5 and f[0] == "traffic" and f[1] == sym:
            return float(f[2]) + float(f[3])
    return None


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """The host threads this process may use: OMP_NUM_THREADS when set, else its CPU affinity."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    if v.isdigit() and int(v) > 0:
        return int(v)
    return len(os.sched_getaffinity(0))


def cpu_baseline(config, log):
    """The oracle (the C restatement of the OPAL 2.0 step: serial-order CIC,
    mixed-radix FFT convolution on the doubled grid, IGF rebuilt every step,
    host libm as the reference binary uses) timed on this host: one full step
    of the benchmark workload itself (1e8 particles, 256^3 grid) on all host
    threads (OpenMP), and the small configuration (1e5 particles, 32^3) on one
    core.  Its bunch is drawn with numpy from the same distribution as the
    GPU's; setup is not timed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes
    import oracle_py as O
    L = O.lib()
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_set_igf_libm.argtypes = [ctypes.c_int]
    L.orc_set_igf_libm(0)
    threads = host_threads()
    model = host_cpu_model()
    out = {"unit": "particle-steps/s", "cores": threads, "kind": "port", "cpu_model": model}

    def bunch(cfg):
        _, bg = gamma_betagamma(cfg["ekin"])
        rng = np.random.default_rng(cfg["seed"])
        N = cfg["N"]
        R = rng.normal(size=(N, 3)) * 1e-3
        if cfg["halo"] > 0:
            m = rng.random(N) < cfg["halo"]
            d = rng.normal(size=(int(m.sum()), 3))
            d /= np.linalg.norm(d, axis=1)[:, None]
            R[m] = d * (5e-3 * rng.random(len(d)) ** (1 / 3))[:, None]
        P = rng.normal(size=(N, 3)) * 1e-3
        P[:, 2] += bg
        return R, P, np.full(N, cfg["q"] / N), np.full(N, 1e-11), bg

    def timed(cfg, steps, nthreads):
        R, P, Q, dt, bg = bunch(cfg)
        n = cfg["n"]
        tr = O.make_tracker((n, n, n), 1e-11, ref_p=(0.0, 0.0, bg), Bz=cfg["Bz"])
        L.orc_set_threads(nthreads)
        t0 = time.perf_counter()
        O.track(tr, R, P, Q, dt, steps)
        t = time.perf_counter() - t0
        return cfg["N"] * steps / t, t

    cfg = CONFIGS[config]
    v, t = timed(cfg, 1, threads)
    log("[bench] cpu: %s, %d threads: one %s step %.1f s" % (model, threads, config, t))
    out["value"] = v
    out["sample"] = ("oracle (CPU restatement of the OPAL 2.0 step, OpenMP, host libm) on %s: one full step "
                     "of the benchmark workload (%d particles, %d^3 grid, %d^3 doubled FFT) on %d threads, %.1f s"
                     % (model, cfg["N"], cfg["n"], 2 * cfg["n"], threads, t))
    v1, t1 = timed(CONFIGS["c1"], 20, 1)
    log("[bench] cpu: c1 on one core: 20 steps %.1f s" % t1)
    out["c1_single_core"] = {"value": v1, "unit": "particle-steps/s", "cores": 1,
                             "sample": "20 steps of 1e5 particles on a 32^3 grid, %.1f s" % t1}
    L.orc_set_threads(1)
    L.orc_set_igf_libm(1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=sorted(CONFIGS))
    ap.add_argument("--sort-interval", type=int, default=20)
    ap.add_argument("--sort-drop", type=int, default=None,
                    help="re-sort by cell blocks: Morton key bits dropped (default: the tuning table's)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-c5", action="store_true", help="skip the energy-binned measurement")
    ap.add_argument("--c5-only", action="store_true",
                    help="run only the energy-binned measurement and print its record (for profiling)")
    ap.add_argument("--no-c4", action="store_true", help="skip the 64-bunch batched measurement")
    ap.add_argument("--no-c2", action="store_true", help="skip the FP64 1e7-particle measurement")
    ap.add_argument("--c4-only", action="store_true",
                    help="run only the 64-bunch batched measurement and print its record (for profiling)")
    ap.add_argument("--c4-fused-only", action="store_true",
                    help="with --c4-only: only the fused batch with the HIP graph")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on exactly one GPU (the space-charge step of one "
              "bunch on one MI355X); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import src_amd as sc
    if args.sort_drop is not None:
        sc.tune_sort_set(args.sort_drop)
    if args.c5_only or args.c4_only:
        log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
        if args.c5_only:
            print(json.dumps(run_c5(sc, args.steps, args.warmup, args.sort_interval, log)))
        else:
            print(json.dumps(run_c4(sc, args.steps, args.warmup, log, args.c4_fused_only)))
        return
    cfg = CONFIGS[args.config]
    N, n = cfg["N"], cfg["n"]
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731

    log("[bench] mixed precision, N=%d n=%d^3" % (N, n))
    cm = make_ctx(sc, cfg, sc.MIXED, args.sort_interval)
    t_mixed, st_mixed = timed_run(cm, args.steps, args.warmup)
    ms_mixed = 1e3 * t_mixed / args.steps
    value = N * args.steps / t_mixed
    log("[bench] mixed: %.3f ms/step, %.3e particle-steps/s" % (ms_mixed, value))

    kernels = {}
    roofline = None
    if not args.no_profile:
        # profile as many steps as the sort interval so the amortised re-sort is included
        ab = algorithmic_bytes(cfg, 4)
        kernels = profile_kernels(cm, max(args.sort_interval, 3), ab)
        dom = max((k for k in kernels if k in ab), key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
        a = kernels[dom]["algorithmic_GBs"]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": a, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                    "traffic": measured_traffic(args.config, dom),
                    "bytes_per_launch": ab[dom], "avg_ms": kernels[dom]["avg_ms"]}
        log("[bench] kernels: " + json.dumps(kernels))
    # achievable bandwidth of a plain copy kernel (also the byte-counter calibration)
    copy_gbs = cm.copy_bandwidth(1 << 30, 20)
    log("[bench] plain copy kernel: %.0f GB/s" % copy_gbs)
    if roofline is not None:
        roofline["copy_kernel_GBs"] = copy_gbs
        roofline["frac_of_copy"] = roofline["achieved"] / copy_gbs

    extra = {}
    if not args.no_fp64:
        log("[bench] FP64 step on the same bunch")
        cd = make_ctx(sc, cfg, sc.FP64, args.sort_interval)
        cm2 = make_ctx(sc, cfg, sc.MIXED, args.sort_interval)
        # one step from the identical bunch: the E-grid difference is the mixed
        # path's error (2^24-scale deposit, FP32 convolution and Grad)
        s64 = cd.track(1)
        s32 = cm2.track(1)
        E64 = cd.get_grid(sc.GRID_EFIELD)
        E32 = cm2.get_grid(sc.GRID_EFIELD)
        e_rel = float(np.linalg.norm(E32 - E64) / np.linalg.norm(E64))
        t64, st64 = timed_run(cd, args.steps, args.warmup)
        _, st32 = timed_run(cm2, args.steps, args.warmup)
        emit_err = max(abs(a["eps_norm"][d] / b["eps_norm"][d] - 1)
                       for a, b in zip(st32, st64) for d in range(3))
        ms64 = 1e3 * t64 / args.steps
        extra = {"fp64_ms_per_step": ms64, "fp64_particle_steps_per_s": N * args.steps / t64,
                 "mixed_over_fp64_speedup": ms64 / ms_mixed, "efield_rel_l2_mixed_vs_fp64": e_rel,
                 "emittance_max_rel_err_mixed_vs_fp64": emit_err,
                 "emittance_steps_compared": len(st64) + 1,
                 "fp64_roofline_step": step_roofline(step_bytes(N, n, 8), ms64)}
        if not args.no_profile:
            k64 = profile_kernels(cd, max(args.sort_interval, 3), algorithmic_bytes(cfg, 8))
            extra["fp64_kernels"] = k64
            g = k64.get("gather_kick_push")
            if g:
                extra["fp64_roofline"] = {"bound": "hbm", "kernel": "gather_kick_push",
                                          "achieved": g["algorithmic_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                          "frac": g["frac_of_peak"], "avg_ms": g["avg_ms"],
                                          "bytes_per_launch": algorithmic_bytes(cfg, 8)["gather_kick_push"]}
        log("[bench] fp64: %.3f ms/step; speedup %.2fx; E rel L2 %.2e; emittance %.2e" %
            (ms64, ms64 / ms_mixed, e_rel, emit_err))
        del s64, s32
        cd.close()
        cm2.close()
    cm.close()

    if not args.no_c5:
        log("[bench] energy-binned step (5e7 particles, 10 bins, 128^3 + image per bin)")
        extra["c5_binned"] = run_c5(sc, args.steps, args.warmup, args.sort_interval, log)
    if not args.no_c4:
        log("[bench] 64 independent bunches, batched")
        extra["c4_batched"] = run_c4(sc, args.steps, args.warmup, log)
    if not args.no_c2:
        # the plain FP64 drop-in configuration: 1e7 particles, 128^3, 1 MeV, Bz 0.5 T
        c2 = CONFIGS["c2"]
        log("[bench] FP64 drop-in step, 1e7 particles, 128^3")
        cc = make_ctx(sc, c2, sc.FP64, args.sort_interval)
        t2, _ = timed_run(cc, args.steps, args.warmup)
        cc.close()
        extra["c2_fp64"] = {"workload": "1e7 macroparticles, 128^3 grid (256^3 doubled FFT), 1 MeV protons, "
                                        "10 nC, Bz 0.5 T, FP64, sort every %d steps" % args.sort_interval,
                            "ms_per_step": 1e3 * t2 / args.steps,
                            "particle_steps_per_s": c2["N"] * args.steps / t2}
        log("[bench] c2 fp64: %.3f ms/step" % extra["c2_fp64"]["ms_per_step"])

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (the oracle on the host cores)")
        cpu = cpu_baseline(args.config, log)
        log("[bench] cpu: %.3e particle-steps/s" % cpu["value"])

    out = {
        "metric": "particle-steps/s @1e8 ptcl, 256^3 grid: mixed vs FP64 at field err<1e-5",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_mixed, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 grid / f64 particles",
        "data": "synthetic (on-device seeded Gaussian bunch)",
        "config": {"workload": "%s: %d macroparticles, %d^3 grid (%d^3 doubled FFT), %g MeV "
                               "protons, %g nC, halo %g, sort every %d steps" %
                               (args.config, N, n, 2 * n, cfg["ekin"], cfg["q"] * 1e9,
                                cfg["halo"], args.sort_interval),
                   "particles": N, "grid": [n, n, n], "precision": "mixed"},
        "roofline": roofline,
        "roofline_step": step_roofline(step_bytes(N, n, 4), ms_mixed),
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
