"""Autotuner of the kernels' tile shapes, run on an MI355X.

Times every candidate on the device and writes src_amd/tuning.json, the table
src_amd/build.py compiles into the library (rebuild after running it):

* FFT tiles: for each transform length L = 2n of the grids the workloads use
  and each precision, every E (elements per thread; a workgroup of 512
  threads holds W = 512 E / L lines of the transform) the kernels take.  The
  measure is a tracker step of a workload of that grid shape: 16 bunches of
  4096 particles solved as one fused batch for n <= 64 (the grid passes of
  the batch dominate), one bunch of 1e6 particles for n = 128 and of 4e6 for
  n = 256.  The fastest E per (L, precision) goes into the table.
* Deposit: the particle count up to which a bunch deposits through the larger
  LDS box (three blocks per CU) rather than the smaller one (four per CU),
  from the deposit's time at 1e6 ... 6.4e7 particles on a 256^3 grid.
* Re-sort granularity: the Hilbert key bits the tracker's re-sort drops (0, 3,
  6: order by cell, by 2^3 or 4^3 blocks of cells), from the step time of a
  2e7-particle bunch on a 256^3 grid re-sorted every 20 steps (40 steps).

Usage: python -m src_amd.autotune [--out src_amd/tuning.json] [--quick]
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import src_amd as sc  # noqa: E402

MP_EV = sc.M_P_EV

# (n, precision, bunches, particles per bunch)
SHAPES = [
    (16, sc.MIXED, 16, 4096), (32, sc.MIXED, 16, 4096), (48, sc.MIXED, 16, 4096), (64, sc.MIXED, 16, 4096),
    (128, sc.MIXED, 1, 1_000_000), (256, sc.MIXED, 1, 4_000_000),
    (16, sc.FP64, 16, 4096), (32, sc.FP64, 16, 4096), (48, sc.FP64, 16, 4096), (64, sc.FP64, 16, 4096),
    (128, sc.FP64, 1, 1_000_000), (256, sc.FP64, 1, 4_000_000),
]
DEPOSIT_SIZES = [1_000_000, 4_000_000, 16_000_000, 64_000_000]


def _ctx(n, precision, N, seed):
    c = sc.SpaceChargeContext((n, n, n), precision=precision, max_particles=N, sort_interval=20)
    bg = 0.7771
    c.generate_bunch(N, seed, (1e-3, 1e-3, 1e-3), (1e-3, 1e-3, 1e-3), bg, 1e-9)
    c.track_init(1e-11, charge=1.0, mass_eV=MP_EV, ref_p=(0.0, 0.0, bg))
    return c


def time_shape(n, precision, bunches, N, steps, reps):
    """ms per tracker step of the shape's workload (best of reps)."""
    cs = [_ctx(n, precision, N, 100 + k) for k in range(bunches)]
    best = math.inf
    try:
        if bunches > 1:
            with sc.Batch(cs) as b:
                b.step(3, want_stats=False)
                for _ in range(reps):
                    t0 = time.perf_counter()
                    b.step(steps, want_stats=False)
                    best = min(best, (time.perf_counter() - t0) / steps)
        else:
            c = cs[0]
            c.track(3, want_stats=False)
            for _ in range(reps):
                t0 = time.perf_counter()
                c.track(steps, want_stats=False)
                best = min(best, (time.perf_counter() - t0) / steps)
    finally:
        for c in cs:
            c.close()
    return 1e3 * best


def tune_fft(log, steps, reps):
    out = []
    for n, prec, bunches, N in SHAPES:
        L = 2 * n
        cands = sc.tune_fft_candidates(L, prec)
        times = {}
        for E in cands:
            sc.tune_fft_set(L, prec, E)
            times[E] = time_shape(n, prec, bunches, N, steps, reps)
        best = min(times, key=times.get)
        sc.tune_fft_set(L, prec, best)  # later shapes of other lengths run with it
        pname = "fp64" if prec == sc.FP64 else "mixed"
        log("[autotune] L=%d %s (%d^3 x %d bunches of %d): %s -> E=%d" %
            (L, pname, n, bunches, N, ", ".join("E=%d %.3f ms" % (e, t) for e, t in times.items()), best))
        out.append({"L": L, "precision": pname, "E": best, "W": best * 512 // L,
                    "workload": "%d^3 grid, %d bunch(es) of %d particles" % (n, bunches, N),
                    "ms_per_step": {str(e): round(t, 4) for e, t in times.items()}})
    return out


def deposit_ms(N, large):
    sc.tune_deposit_set(N if large else 0)
    c = _ctx(256, sc.MIXED, N, 7)
    try:
        c.track(2, want_stats=False)
        c.profile_enable(True)
        c.track(10)
        prof = c.profile_read()
        c.profile_enable(False)
    finally:
        c.close()
    ms, k = prof["deposit"]
    return ms / k


def tune_deposit(log):
    res = {}
    for N in DEPOSIT_SIZES:
        res[N] = (deposit_ms(N, True), deposit_ms(N, False))
        log("[autotune] deposit %d particles: large box %.4f ms, small box %.4f ms" % (N, res[N][0], res[N][1]))
    # the largest size at which the larger box is still faster; the threshold
    # lies halfway (geometrically) to the next size
    limit = 0
    for i, N in enumerate(DEPOSIT_SIZES):
        if res[N][0] <= res[N][1]:
            nxt = DEPOSIT_SIZES[i + 1] if i + 1 < len(DEPOSIT_SIZES) else 4 * N
            limit = int(math.sqrt(N * nxt))
        else:
            break
    return {"large_box_max_particles": limit,
            "ms": {str(N): {"large": round(a, 4), "small": round(b, 4)} for N, (a, b) in res.items()}}


def tune_sort(log, reps):
    res = {}
    for drop in (0, 3, 6):
        sc.tune_sort_set(drop)
        c = _ctx(256, sc.MIXED, 20_000_000, 9)
        try:
            c.track(3, want_stats=False)
            best = math.inf
            for _ in range(reps):
                t0 = time.perf_counter()
                c.track(40, want_stats=False)
                best = min(best, (time.perf_counter() - t0) / 40)
        finally:
            c.close()
        res[drop] = 1e3 * best
        log("[autotune] re-sort dropping %d key bits: %.4f ms/step" % (drop, res[drop]))
    best = min(res, key=res.get)
    sc.tune_sort_set(best)
    return {"drop_bits": best, "ms_per_step": {str(k): round(v, 4) for k, v in res.items()},
            "workload": "256^3 grid, 2e7 particles, re-sort every 20 steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.time()
    for n, prec, _, _ in SHAPES:  # start from the default rule, not an earlier table
        sc.tune_fft_set(2 * n, prec, 0)
    table = {"generated_by": "python -m src_amd.autotune", "fft_tiles": tune_fft(log, args.steps, args.reps),
             "deposit": tune_deposit(log), "sort": tune_sort(log, args.reps)}
    table["seconds"] = round(time.time() - t0, 1)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
        f.write("\n")
    log("[autotune] wrote %s in %.0f s" % (args.out, table["seconds"]))


if __name__ == "__main__":
    main()
