"""Benchmark of the time-lagged pair-correlation path on one MI355X.

Default workload (the measured line, --config c3): 65,536 atoms of 2 types
(50/50), 1,000 frames of a seeded Gaussian random walk (sigma 0.08 A per
axis) in a cubic 90 A cell; Gofrt in the FP32 + FP64-recheck form with rmin
0, rmax 15 A, 512 bins, lags 0..64, t0 stride 1, reset(936), calculate(0).
One step = one whole Gofrt::calculate.  The reference evaluates every ordered
atom pair for every (t0, lag), 65536^2 * 936 * 65 = 2.61e14 pair-lag updates
per step; this build screens pairs and prunes whole groups of far pairs, so
the reported value counts only the pairs some distance test evaluated (a
device counter), and the all-pairs rate is reported beside it as the
time-to-solution rate.  Both arithmetic forms are timed on one t0 slice of
the same workload and their counts compared.  Inputs are resident in HBM
before timing starts; the launch shape is autotuned (untimed) first.  The
same line carries the MSD of that trajectory (lags 0..498 with the
centre-of-mass channel and drift removal, reset(499)) and the Green-Kubo
workload (c5) as sub-records.  The CPU baseline is the oracle's restatement
of the reference loop nest on the host cores, on a bounded t0 slice.

Other workloads (--config), each printing one JSON line of its own:
  c1  1,000 atoms 500/500, cubic 32.2 A, 200 frames: Gofrt rmax 10, 200 bins,
      lags 0..50, reset(150); the CPU baseline runs it whole
  c2  16,384 atoms of 3 types (80/10/10 %), 2,000 frames, sigma 0.05 A,
      triclinic cell with diagonal 56 A and tilts uniform in +-0.1 L (fixed
      per run); Gofrt rmax 12 A, 256 bins, lags 0..100, reset(1900)
  c4  8,192 atoms 80/10/10, cubic 45 A, 60,000 frames; 64 seeded windows of
      500..4000 frames: Gofrt lags 0..40 and MSD lags 0..400 of every window in
      one graph launch, with the mean and variance over windows
  c5  Green-Kubo, 2 vector currents (6 FP64 columns), 2,000,000 steps of
      seeded AR(2) processes with cross-coupling 0.3, 32 uneven blocks
      partitioning t0 in [0, 1,980,000), 20,000 lags per block, lag products
      on the matrix cores (split-BF16 MFMA) vs the FP64 form
  sh  SphericalCorrelations (lmax 10): 2,048 atoms 50/50 at 0.1 atoms/A^3
      (cubic 27.4 A), 2,200 frames, sigma 0.05 A; ranges (0.7, 3.5) A for every
      type pair, 4 bins, lags 0..199, t0 stride 10, reset(2000), calculate(0);
      also the Steinhardt (q4, q6) histogram of every frame of that trajectory
  lammps  g(r,t) block average (6 blocks, lags 0..40, 256 bins to 12 A) over a
      LAMMPS-binary file of 8,192 atoms x 600 frames read in windows, with the
      next block's window read ahead during the current block, and without

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--config c3]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_UPDATE = {True: 19, False: 16}   # triclinic / orthorhombic pair-lag update
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
MSD_FLOP = 9                               # per atom-lag-t0 update, +3 with drift removal


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class Hip:
    """The few HIP runtime calls the timing needs (events on the launch stream)."""

    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")

    def _c(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed with HIP error %d" % (what, rc))

    def stream(self):
        s = ctypes.c_void_p()
        self._c(self.h.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self._c(self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 8))), "hipMalloc")
        return p

    def free(self, p):
        self.h.hipFree(p)

    def event(self):
        e = ctypes.c_void_p()
        self._c(self.h.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e

    def record(self, e, s):
        self._c(self.h.hipEventRecord(e, s), "hipEventRecord")

    def elapsed_ms(self, a, b):
        self._c(self.h.hipEventSynchronize(b), "hipEventSynchronize")
        ms = ctypes.c_float()
        self._c(self.h.hipEventElapsedTime(ctypes.byref(ms), a, b), "hipEventElapsedTime")
        return ms.value

    def sync(self):
        self._c(self.h.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def d2h(self, host, dev, nbytes):
        self._c(self.h.hipMemcpy(ctypes.c_void_p(host.ctypes.data), dev, ctypes.c_size_t(nbytes), 2), "hipMemcpy")


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

def gofrt_workload(name):
    from analisi_amd import synth
    if name == "c1":
        seed, N, F, L, fr, sigma, tilt = 11, 1000, 200, 32.2, [0.5, 0.5], 0.05, None
        g = dict(rmin=0.0, rmax=10.0, nbin=200, tmax=51, n=150)
    elif name == "c2":
        seed, N, F, L, fr, sigma = 2024, 16384, 2000, 56.0, [0.8, 0.1, 0.1], 0.05
        tilt = (2.0 * synth.uniform(seed, 7, 3) - 1.0) * 0.1 * L
        g = dict(rmin=0.0, rmax=12.0, nbin=256, tmax=101, n=1900)
    elif name == "c3":
        seed, N, F, L, fr, sigma, tilt = 33, 65536, 1000, 90.0, [0.5, 0.5], 0.08, None
        g = dict(rmin=0.0, rmax=15.0, nbin=512, tmax=65, n=936, msd_tmax=501, msd_n=499)
    elif name == "c4":
        seed, N, F, L, fr, sigma, tilt = 44, 8192, 60000, 45.0, [0.8, 0.1, 0.1], 0.05, None
        g = dict(rmin=0.0, rmax=12.0, nbin=256, tmax=41, n=0, msd_tmax=401, nblocks=64)
    else:
        raise ValueError(name)
    box_l = synth.lammps_box(L, tilt)
    log("generating %s: N=%d F=%d" % (name, N, F))
    t = time.perf_counter()
    pos = synth.random_walk(seed, N, F, box_l, sigma,
                            progress=lambda f: log("generated %d of %d frames (%.0f s)" % (f, F, time.perf_counter() - t)))
    log("generated in %.1f s" % (time.perf_counter() - t))
    types = synth.types_by_fraction(N, fr)
    w = dict(name=name, pos=pos, types=types, box_l=box_l, tilt=tilt, N=N, F=F, L=L, sigma=sigma, seed=seed,
             triclinic=tilt is not None, skip=1, every=1, primo=0, fractions=fr)
    w.update(g)
    return w


def describe(w):
    cell = "triclinic, diagonal %g A, tilts %s" % (w["L"], np.round(w["tilt"], 4).tolist()) if w["triclinic"] \
        else "cubic %g A" % w["L"]
    return ("%d atoms %s, %d frames, %s, sigma %g A (seed %d)"
            % (w["N"], "/".join("%g" % (100 * f) for f in w["fractions"]), w["F"], cell, w["sigma"], w["seed"]))


def internal_box(box_l, F):
    """LAMMPS box -> the internal [lo(3), half-length(3), tilt(3)] format."""
    b = np.array(box_l, dtype=np.float64).copy()
    if b.size == 6:
        out = np.array([b[0], b[2], b[4], (b[1] - b[0]) / 2, (b[3] - b[2]) / 2, (b[5] - b[4]) / 2])
    else:
        out = np.array([b[0], b[2], b[4], (b[1] - b[0]) / 2, (b[3] - b[2]) / 2, (b[5] - b[4]) / 2,
                        b[6], b[7], b[8]])
    return np.tile(out, (F, 1))


def type_ids(types):
    """Dense type ids in sorted order of the type values (basetrajectory.cpp:51-89)."""
    u, inv = np.unique(np.asarray(types), return_inverse=True)
    return inv.astype(np.int32), len(u)


def centres_of_mass(pos, tid, nt):
    """Per-type centre of mass per frame, unit masses, [F][nt][3], as the
    reference's running mean over atoms in index order (cm += (x - cm)/n,
    trajectory_numpy.cpp:201-225), vectorised over frames."""
    F = pos.shape[0]
    cm = np.zeros((F, nt, 3))
    cont = np.zeros(nt, dtype=np.int64)
    for ia in range(pos.shape[1]):
        t = tid[ia]
        cont[t] += 1
        cm[:, t, :] += (pos[:, ia, :] - cm[:, t, :]) / float(cont[t])
    return cm


def self_channel(nt, t):
    """Channel of the (t, t) self part: get_itype (gofrt.h:86-104) + nt(nt+1)/2."""
    half = nt * (nt + 1) // 2
    return half + half - (t + 1) * (t + 2) // 2 + t


def n_updates(N, n, tmax, skip=1, every=1):
    """Ordered pairs x t0 values x lags of Gofrt::calculate after reset(n):
    leff = min(n, tmax), lags 0, every, .. < leff, t0 = 0, skip, .. < n."""
    leff = n if (n < tmax or tmax == 0) else tmax
    return float(N) * N * ((n + skip - 1) // skip) * ((leff + every - 1) // every)


def device_traj(w, with_cm=False):
    from analisi_amd import capi
    tid, nt = type_ids(w["types"])
    cm = centres_of_mass(w["pos"], tid, nt) if with_cm else None
    boxi = internal_box(w["box_l"], w["F"])
    return capi.DeviceTrajectory(w["pos"], boxi, tid, nt, cm), tid, nt


def cpu_gofrt(w, n, nthreads, F_used=None):
    """The oracle (reference loop nest) on reset(n) of the same trajectory."""
    from oracle import oracle as O
    F = F_used or (min(n, w["tmax"]) + n)
    tr = O.Traj(w["pos"][:F], w["types"], np.tile(w["box_l"], (F, 1)), w["triclinic"])
    t = time.perf_counter()
    counts, _ = O.gofrt(tr, w["rmin"], w["rmax"], w["nbin"], w["tmax"], nthreads, 1, 1, n, 0, want_vdata=True)
    dt = time.perf_counter() - t
    return n_updates(w["N"], n, w["tmax"]) / dt, dt, counts


def traffic_from_profile(kernel_prefix, workload):
    """HBM bytes per launch of a kernel, from a committed rocprofv3 counter
    summary under profiles/ (lines 'kernel workload bytes_per_launch')."""
    p = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and not line.startswith("#") and parts[0].startswith(kernel_prefix) \
                    and parts[1] == workload:
                return float(parts[2])
    return None


# ---------------------------------------------------------------------------
# g(r,t) (c1, c2, c3)
# ---------------------------------------------------------------------------

def profile_figure(key, workload):
    """A figure measured by rocprofv3 counters and committed under profiles/
    (lines 'key workload value' in profiles/counters.txt), or None."""
    p = os.path.join(ROOT, "profiles", "counters.txt")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and not line.startswith("#") and parts[0] == key and parts[1] == workload:
                return float(parts[2])
    return None


def host_cores():
    """Threads the CPU baseline uses (at most 16, one GPU's share of the host's cores)
    and what the host reports."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = nproc
    return min(aff, 16), nproc, aff


def timed_calculates(L, g, w, d_out, stream, hip, nsteps, what):
    """nsteps whole Gofrt::calculate calls on one stream, HIP events around
    each; every call first waits for the device (it uploads its work list),
    so a step's time includes the host's launch work.  A line per step."""
    check = __import__("analisi_amd.capi", fromlist=["check"]).check
    evs = [hip.event() for _ in range(nsteps + 1)]
    hip.sync()
    t_wall = time.perf_counter()
    hip.record(evs[0], stream)
    for i in range(nsteps):
        check(L.anl_gofrt_calculate_async(g.handle, w["primo"], d_out, stream))
        hip.record(evs[i + 1], stream)
        if i > 0:
            log("%s %d: %.1f ms" % (what, i - 1, hip.elapsed_ms(evs[i - 1], evs[i])))
    hip.sync()
    t_wall = time.perf_counter() - t_wall
    step_ms = [hip.elapsed_ms(evs[i], evs[i + 1]) for i in range(nsteps)]
    if nsteps:
        log("%s %d: %.1f ms" % (what, nsteps - 1, step_ms[-1]))
    return step_ms, t_wall


def run_gofrt(w, args, hip):
    from analisi_amd import capi
    check = capi.check
    L = capi.lib()
    t = time.perf_counter()
    traj, tid, nt = device_traj(w, with_cm=("msd_tmax" in w))
    log("device trajectory ready in %.1f s" % (time.perf_counter() - t))
    N, F = w["N"], w["F"]
    g = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], w["skip"], w["every"], capi.GOFRT_SCREENED)
    leff = g.reset(w["n"])
    tune = None
    if args.shape:
        u, x, lp, cl = (int(v) for v in args.shape.split(","))
        g.set_tuning(u, x, 1 << lp, cl)
        tune = {"given": [u, x, 1 << lp, cl]}
    elif not args.no_autotune:
        nl = max(1, min(leff, args.tune_lags))
        ms, best = g.autotune(w["primo"], nl)
        tune = {"probe_lags": nl, "candidates_ms": [round(float(x), 2) for x in ms], "best": int(best),
                "chosen": list(g.tuning())}
        log("autotune on %d lags: %s -> %s" % (nl, tune["candidates_ms"], tune["chosen"]))
    nch = nt * (nt + 1)
    out_len = leff * nch * w["nbin"]
    d_out = hip.malloc(out_len * 8)
    stream = hip.stream()
    updates = n_updates(N, w["n"], w["tmax"], w["skip"], w["every"])
    timed_calculates(L, g, w, d_out, stream, hip, args.warmup, "warmup")
    step_ms, t_wall = timed_calculates(L, g, w, d_out, stream, hip, args.steps, "step")
    ms = float(np.mean(step_ms)) if step_ms else float("nan")
    counts = np.zeros((leff, nch, w["nbin"]), dtype=np.uint64)
    hip.d2h(counts, d_out, out_len * 8)
    _, ncand, nre = g.stats()
    _, nexact = g.recheck_stats()
    ps = g.pair_stats()
    clustered, units = g.cluster_stats()
    # invariants of the full run: lag-0 self channels hold N_type per t0 in
    # bin 0, the pair channels are symmetric (even counts at lag 0)
    ntype = np.bincount(tid, minlength=nt)
    half = nt * (nt + 1) // 2
    inv_ok = True
    for t_ in range(nt):
        ch = self_channel(nt, t_)
        inv_ok &= int(counts[0, ch, 0]) == int(ntype[t_]) * w["n"] and int(counts[0, ch, 1:].sum()) == 0
    inv_ok &= bool(np.all(counts[0, :half] % 2 == 0))

    # both forms on one t0 slice of the same workload, timed alike; counts compared
    n_sl = min(args.fp64_n or (101 if w["name"] == "c2" else 16), w["n"])
    ws = dict(w, n=n_sl)
    g64 = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_FP64)
    g64.reset(n_sl)
    gsc = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_SCREENED)
    gsc.set_tuning(*g.tuning())
    gsc.reset(n_sl)
    leff_sl = g64.leff
    len_sl = leff_sl * nch * w["nbin"]
    d64, dsc = hip.malloc(len_sl * 8), hip.malloc(len_sl * 8)
    timed_calculates(L, g64, ws, d64, stream, hip, 1, "fp64 slice warmup")
    ms64 = timed_calculates(L, g64, ws, d64, stream, hip, 1, "fp64 slice")[0][0]
    timed_calculates(L, gsc, ws, dsc, stream, hip, 1, "fp32+recheck slice warmup")
    mssl = timed_calculates(L, gsc, ws, dsc, stream, hip, 1, "fp32+recheck slice")[0][0]
    c64 = np.zeros((leff_sl, nch, w["nbin"]), dtype=np.uint64)
    csc = np.zeros_like(c64)
    hip.d2h(c64, d64, len_sl * 8)
    hip.d2h(csc, dsc, len_sl * 8)
    slice_identical = bool(np.array_equal(c64, csc))
    upd64 = n_updates(N, n_sl, w["tmax"])
    rate64 = upd64 / (ms64 * 1e-3)
    log("slice reset(%d): fp64 %.1f ms, fp32+recheck %.1f ms, identical=%s" % (n_sl, ms64, mssl, slice_identical))

    evaluated = ps["evaluated"]
    fl = FLOP_PER_UPDATE[w["triclinic"]]
    sec = ms * 1e-3
    achieved = evaluated * fl / sec / 1e12
    cname = w["name"]
    busy = profile_figure("gofrt_screened_valu_busy", cname)
    res = {
        "metric": "g(r,t) pair-lag updates/s, FP32+recheck vs plain FP64, bit-exact counts",
        "value": evaluated / sec,
        "unit": "pair-lag updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i16/f16 screens, f32 candidates, f64 recheck",
        "data": "synthetic (seeded SplitMix64/Box-Muller random walk, seed %d)" % w["seed"],
        "config": {"workload": "%s: Gofrt g(r,t) FP32+recheck, %s; rmax %g A, %d bins, lags 0..%d, t0 stride 1, "
                               "reset(%d), calculate(0)" % (cname.upper(), describe(w), w["rmax"], w["nbin"],
                                                            leff - 1, w["n"]),
                   "form": "fp32+recheck", "natoms": N, "frames": F, "lags": leff, "t0": w["n"], "nbin": w["nbin"],
                   "launch_shape": list(g.tuning())},
        "pairs": {"evaluated_per_step": int(evaluated), "all_per_step": int(ps["all"]),
                  "reference_count_per_step": updates, "tile_screen": int(ps["tile"]),
                  "screen16": int(ps["screen16"]), "pruned_fraction": 1.0 - evaluated / max(ps["all"], 1),
                  "note": "value counts evaluated pairs only: a pair is evaluated when a distance test looks at that "
                          "very pair (FP16 MFMA tile, 16-bit screen, FP32/FP64 evaluation); pairs of 32-column halves "
                          "skipped whole by the cluster bounding-sphere test are pruned and not counted"},
        "reference_equivalent": {"value": updates / sec, "unit": "pair-lag updates/s",
                                 "note": "all N^2 n0 Lf ordered pairs the reference evaluates, per second of whole "
                                         "calculate (time-to-solution rate, not work done)"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": traffic_from_profile("gofrt_screened", cname),
                     "peak_nonpacked_fp32": FP32_PEAK_TFLOPS / 2,
                     "frac_nonpacked": achieved / (FP32_PEAK_TFLOPS / 2),
                     "valu_busy_measured": busy,
                     "note": "%d algorithmic FLOP per evaluated pair-lag update (3 sub, 6 min image, %s5 d2, 2 gates) "
                             "x evaluated pairs / event-timed whole calculate; peak = MI355X published FP32 157.3 "
                             "TFLOP/s (packed); this build issues no packed FP32, whose half, 78.6, is the "
                             "non-packed VALU figure.  The kernel is instruction-issue bound (VALU/SALU/LDS), so "
                             "neither 'hbm' nor 'mfma' describes it; valu_busy_measured = SQ_INSTS_VALU x 4 / (1024 "
                             "SIMDs x SQ_BUSY_CYCLES / 32), from profiles/counters.txt" % (fl, "3 tilt, " if w["triclinic"] else "")},
        "time_to_solution_vs_fp64": {
            "slice": "reset(%d) of the same trajectory: %d t0 x %d lags, both forms timed alike" % (n_sl, n_sl, leff_sl),
            "fp64_ms": ms64, "fp32_recheck_ms": mssl, "ratio_on_slice": ms64 / mssl,
            "fp64_rate": rate64, "ratio_whole_run": (updates / sec) / rate64,
            "counts_identical_on_slice": slice_identical,
            "fp64_roofline_frac": upd64 * fl / (ms64 * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
            "note": "ratio_whole_run = all-pairs rate of the whole FP32+recheck calculate over the FP64 form's "
                    "all-pairs rate on the slice (the FP64 form evaluates every pair; its rate does not depend "
                    "on t0 or lag)"},
        "screen": {"candidates_per_step": int(ncand), "candidate_fraction_of_all": ncand / updates,
                   "fp64_rechecks_per_step": int(nre), "recheck_fraction_of_candidates": nre / max(ncand, 1),
                   "exact_rechecks_per_step": int(nexact), "exact_fraction_of_rechecks": nexact / max(nre, 1),
                   "note": "rechecks: candidates whose FP32 distance fell in the window around an edge, "
                           "re-evaluated in FP64 from their 32-bit fractions; exact: those still within that "
                           "evaluation's error bound of an edge, evaluated with the reference's FP64 sequence "
                           "from the positions",
                   "clustered_units": int(clustered), "units": int(units)},
        "checks": {"lag0_self_channels_and_pair_symmetry": bool(inv_ok), "all_pairs_equal_reference_count":
                   int(ps["all"]) == int(updates)},
        "autotune": tune,
        "wall_s_timed": t_wall,
    }
    if not args.no_cpu:
        nthr, nproc, aff = host_cores()
        cn = w["n"] if cname == "c1" else (args.cpu_n or (2 if cname == "c3" else 8))
        log("cpu baseline: oracle on reset(%d), %d threads" % (cn, nthr))
        cr, cdt, ccounts = cpu_gofrt(w, cn, nthr, F_used=F if cname == "c1" else None)
        if cn == w["n"]:
            same = bool(np.array_equal(counts, ccounts))
        else:   # the oracle sample must equal the device result on the same reset(n)
            gsl = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_SCREENED)
            gsl.reset(cn)
            same = bool(np.array_equal(gsl.calculate(0), ccounts))
        res["cpu_baseline"] = {"value": cr, "unit": "pair-lag updates/s", "cores": nthr, "kind": "port",
                               "host_cpus": nproc, "affinity_cpus": aff,
                               "sample": "oracle (reference loop nest, one thread round per (lag, t0), every ordered "
                                         "pair) on the same trajectory and histogram, reset(%d): %d t0 x %d lags, "
                                         "%.1f s on %d threads" % (cn, cn, min(cn, w["tmax"]), cdt, nthr),
                               "counts_identical_to_gpu": same}
        log("cpu baseline %.3e upd/s on %d threads (%.1f s), identical=%s" % (cr, nthr, cdt, same))
    for p_ in (d_out, d64, dsc):
        hip.free(p_)
    if "msd_tmax" in w:
        res["msd"] = run_msd(w, args, hip, traj, nt)
    return res


def run_msd(w, args, hip, traj, nt):
    """MSD with the centre-of-mass channel and drift removal, timed on the
    device; the FP64 oracle error on a t0 slice."""
    from analisi_amd import capi
    check = capi.check
    L = capi.lib()
    n, tmax = w["msd_n"], w["msd_tmax"]
    leff = n if n < tmax else tmax
    out_len = leff * 2 * nt
    d = hip.malloc(out_len * 8)
    stream = hip.stream()
    tune = None
    if not args.no_autotune:
        ms_t, best = traj.msd_autotune(1, tmax, True, True, n, 0)
        tune = {"shapes": [list(x) for x in traj.MSD_SHAPES], "candidates_ms": [round(float(x), 3) for x in ms_t],
                "chosen_shape": list(traj.msd_shape()),
                "shape": "lag tile, wavefronts per workgroup (tile 0: eight lags per thread)"}
        log("msd autotune: %s -> shape %s" % (tune["candidates_ms"], traj.msd_shape()))
    check(L.anl_msd_calculate_async(traj.handle, 1, tmax, 1, 1, n, 0, d, stream))
    hip.sync()
    e0, e1 = hip.event(), hip.event()
    reps = max(1, args.steps)
    hip.record(e0, stream)
    for _ in range(reps):
        check(L.anl_msd_calculate_async(traj.handle, 1, tmax, 1, 1, n, 0, d, stream))
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1) / reps
    upd = float(w["N"]) * n * leff
    flop = MSD_FLOP + 3
    res = {"metric": "MSD atom-lag-t0 updates/s", "value": upd / (ms * 1e-3), "ms_per_step": ms,
           "config": "MSD lags 0..%d, reset(%d), centre-of-mass channel + drift removal, skip 1" % (leff - 1, n),
           "autotune": tune,
           "roofline": {"bound": "valu", "achieved": upd * flop / (ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": upd * flop / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                        "note": "%d FLOP per atom-lag-t0 update (3 sub, 3 mul, 2 add, 1 accumulate, "
                                "3 drift)" % flop}}
    hip.free(d)
    if not args.no_cpu:
        from oracle import oracle as O
        ns = args.msd_cpu_n
        F = ns + min(ns, tmax)
        tr = O.Traj(w["pos"][:F], w["types"], np.tile(w["box_l"], (F, 1)), w["triclinic"])
        nthr, nproc, aff = host_cores()
        t = time.perf_counter()
        ref = O.msd(tr, 1, tmax, nthr, True, True, ns, 0)
        dt = time.perf_counter() - t
        got = capi.msd(traj, 1, tmax, True, True, ns, 0)
        err = float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)))
        res["cpu_baseline"] = {"value": float(w["N"]) * ns * min(ns, tmax) / dt, "unit": "atom-lag-t0 updates/s",
                               "cores": nthr, "kind": "port", "host_cpus": nproc, "affinity_cpus": aff,
                               "sample": "oracle on the same trajectory, reset(%d): %d t0 x %d lags, %.1f s"
                                         % (ns, ns, min(ns, tmax), dt),
                               "max_rel_error_vs_fp64_oracle": err}
        log("msd: %.1f ms, oracle slice max rel error %.2e" % (ms, err))
    return res


# ---------------------------------------------------------------------------
# block engine (c4)
# ---------------------------------------------------------------------------

def c4_windows(w, nb=None):
    from analisi_amd import synth
    nb, F = nb or w["nblocks"], w["F"]
    u1 = synth.uniform(w["seed"], 21, nb)
    u2 = synth.uniform(w["seed"], 22, nb)
    lens = (500 + np.floor(u1 * 3501)).astype(np.int64).clip(500, 4000)
    starts = np.floor(u2 * (F - lens - 401)).astype(np.int64)
    return starts.astype(np.uint64), lens.astype(np.uint32)


def run_blocks(w, args, hip):
    """The block engine on the c4 windows.  With --c4-split K the windows are
    cut into K consecutive groups, one engine (one graph) each, replayed one
    after the other inside the timed step; K = 1 is one graph for all."""
    from analisi_amd import capi
    traj, tid, nt = device_traj(w, with_cm=False)
    starts, lens = c4_windows(w, args.c4_windows)
    g = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_SCREENED)
    tune = None
    if not args.no_autotune:
        g.reset(int(lens.min()))
        ms_c, best = g.autotune(int(starts[0]), 2)
        tune = {"candidates_ms": [round(float(x), 2) for x in ms_c], "chosen": list(g.tuning())}
        log("autotune: %s -> %s" % (tune["candidates_ms"], tune["chosen"]))
    K = max(1, args.c4_split)
    cuts = np.linspace(0, len(lens), K + 1).astype(int)
    t = time.perf_counter()
    engines = [capi.DeviceBlocks(traj, starts[a:b], lens[a:b], gofrt=g, msd=True, msd_skip=1,
                                 msd_tmax=w["msd_tmax"], use_graph=True) for a, b in zip(cuts[:-1], cuts[1:])]
    t_create = time.perf_counter() - t
    log("block engine ready: %d windows in %d graph(s), %d frames in all, set up in %.1f s"
        % (len(lens), K, int(lens.sum()), t_create))
    stream = hip.stream()
    for i in range(args.warmup):
        for B in engines:
            B.run(stream)
        hip.sync()
        log("warmup %d done" % i)
    reps = max(1, args.steps)
    ms = 0.0
    for r in range(reps):
        for k, B in enumerate(engines):
            e0, e1 = hip.event(), hip.event()
            hip.record(e0, stream)
            B.run(stream)
            hip.record(e1, stream)
            dt = hip.elapsed_ms(e0, e1)
            ms += dt
            log("step %d graph %d/%d: %.1f ms" % (r, k + 1, K, dt))
    ms /= reps
    N = w["N"]
    upd = sum(n_updates(N, int(n), w["tmax"]) for n in lens)
    # every window's g(r,t) runs in one launch, so the handle's pair counters
    # cover the whole step
    ps = g.pair_stats()
    evaluated = int(ps["evaluated"])
    mupd = sum(float(N) * int(n) * min(int(n), w["msd_tmax"]) for n in lens)
    # one window recomputed alone: FP64 g(r,t) and MSD, bit-identical
    b = int(np.argmin(lens))
    k = int(np.searchsorted(cuts, b, side="right") - 1)
    gmean, gvar, gcounts = engines[k].gofrt_results()
    mmean, mvar, mper = engines[k].msd_results()
    bl = b - cuts[k]
    g64 = capi.DeviceGofrt(traj, w["rmin"], w["rmax"], w["nbin"], w["tmax"], 1, 1, capi.GOFRT_FP64)
    g64.reset(int(lens[b]))
    same_g = bool(np.array_equal(g64.calculate(int(starts[b])), gcounts[bl]))
    # the plain FP64 form timed on that window (the comparator of the batched run)
    nch = nt * (nt + 1)
    leff_b = min(int(lens[b]), w["tmax"])
    d64 = hip.malloc(leff_b * nch * w["nbin"] * 8)
    wb = dict(w, primo=int(starts[b]), n=int(lens[b]))
    ms64 = timed_calculates(capi.lib(), g64, wb, d64, stream, hip, 1, "fp64 window %d" % b)[0][0]
    hip.free(d64)
    rate64 = n_updates(w["N"], int(lens[b]), w["tmax"]) / (ms64 * 1e-3)
    m1 = capi.msd(traj, 1, w["msd_tmax"], False, False, int(lens[b]), int(starts[b]))
    same_m = bool(np.array_equal(m1, mper[bl]))
    res = {"metric": "block-averaged g(r,t) + MSD, evaluated pair-lag updates/s over all windows",
           "value": evaluated / (ms * 1e-3), "unit": "pair-lag updates/s", "n_gpus": 1, "ms_per_step": ms,
           "steps": reps, "dtype": "i16 screen + f32 + f64 recheck; MSD f64",
           "config": {"workload": "block engine: %s; %d windows of 500..4000 frames; Gofrt rmax %g, %d bins, "
                                  "lags 0..%d; MSD lags 0..%d; %d HIP graph(s) per step"
                                  % (describe(w), len(lens), w["rmax"], w["nbin"], w["tmax"] - 1,
                                     w["msd_tmax"] - 1, K),
                      "sum_window_frames": int(lens.sum()), "gofrt_updates": upd, "msd_updates": mupd,
                      "launch_shape": list(g.tuning())},
           "pairs": {"evaluated_per_step": evaluated, "all_per_step": int(ps["all"]),
                     "reference_count_per_step": upd, "pruned_fraction": 1.0 - evaluated / max(int(ps["all"]), 1),
                     "note": "value counts the pairs some distance test evaluated (FP16 MFMA tile, 16-bit screen, "
                             "FP32/FP64 evaluation) in the step's one g(r,t) launch; groups skipped whole are pruned"},
           "reference_equivalent": {"value": upd / (ms * 1e-3), "unit": "pair-lag updates/s",
                                    "note": "all N^2 n0 Lf ordered pairs of every window per second of the whole step "
                                            "(g(r,t) + MSD of all windows)"},
           "setup_s": t_create, "autotune": tune,
           "time_to_solution_vs_fp64": {"window": b, "frames": int(lens[b]), "fp64_ms": ms64, "fp64_rate": rate64,
                                        "ratio_whole_run": upd / (ms * 1e-3) / rate64,
                                        "note": "plain FP64 form timed on one window's reset(n) calculate; ratio = "
                                                "all-pairs rate of the whole batched g(r,t) + MSD step over that"},
           "checks": {"window_%d_gofrt_fp64_identical" % b: same_g, "window_%d_msd_identical" % b: same_m,
                      "all_pairs_equal_reference_count": int(ps["all"]) == int(upd)}}
    if not args.no_cpu:
        from oracle import oracle as O
        nw = cuts[k + 1] - cuts[k]
        mv = O.MediaVar(gmean.size)
        for q in range(nw):
            mv.add(gcounts[q].ravel().astype(np.float64) * (1.0 / int(lens[cuts[k] + q])))
        mv.end(nw)
        mv2 = O.MediaVar(mmean.size)
        for q in range(nw):
            mv2.add(mper[q].ravel())
        mv2.end(nw)
        # the oracle's g(r,t) on a bounded slice of the same window, timed on
        # the host cores, its counts checked against the device's
        nthr, nproc, aff = host_cores()
        cn = args.cpu_n or 16
        Fs = min(cn, w["tmax"]) + cn
        sl = w["pos"][int(starts[b]):int(starts[b]) + Fs]
        ot = O.Traj(sl, w["types"], np.tile(w["box_l"], (Fs, 1)), w["triclinic"])
        log("cpu baseline: oracle on window %d, reset(%d), %d threads" % (b, cn, nthr))
        t = time.perf_counter()
        cc, _ = O.gofrt(ot, w["rmin"], w["rmax"], w["nbin"], w["tmax"], nthr, 1, 1, cn, 0, want_vdata=False)
        cdt = time.perf_counter() - t
        g64.reset(cn)
        same_cpu = bool(np.array_equal(g64.calculate(int(starts[b])), cc))
        res["cpu_baseline"] = {"value": n_updates(w["N"], cn, w["tmax"]) / cdt, "unit": "pair-lag updates/s",
                               "cores": nthr, "kind": "port", "host_cpus": nproc, "affinity_cpus": aff,
                               "sample": "oracle (reference loop nest, every ordered pair) on window %d, reset(%d): "
                                         "%d t0 x %d lags, %.1f s on %d threads" % (b, cn, cn, min(cn, w["tmax"]),
                                                                                   cdt, nthr),
                               "counts_identical_to_gpu": same_cpu}
        log("cpu baseline %.3e upd/s on %d threads (%.1f s), identical=%s"
            % (res["cpu_baseline"]["value"], nthr, cdt, same_cpu))
        res["checks"]["mediavar_identical_to_oracle"] = bool(
            np.array_equal(gmean.ravel(), mv.mean) and np.array_equal(gvar.ravel(), mv.var)
            and np.array_equal(mmean.ravel(), mv2.mean) and np.array_equal(mvar.ravel(), mv2.var))
    log("blocks: %.1f ms/step, checks %s" % (ms, res["checks"]))
    for B in engines:
        B.close()
    return res


# ---------------------------------------------------------------------------
# Green-Kubo (c5)
# ---------------------------------------------------------------------------

def c5_blocks(seed, T, lmax, nb):
    from analisi_amd import synth
    span = T - lmax
    u = synth.uniform(seed, 31, nb)
    extra = span - nb * lmax
    lens = lmax + np.floor(extra * u / u.sum()).astype(np.int64)
    lens[-1] += span - lens.sum()
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    return starts, lens


def run_gk(args, hip):
    from analisi_amd import capi, synth
    check = capi.check
    L = capi.lib()
    T, lmax, nb, seed = 2_000_000, 20_000, 32, 55
    log("generating AR(2) currents, %d steps" % T)
    J = synth.ar2_currents(seed, T, ncur=2, coupling=0.3)
    starts, lens = c5_blocks(seed, T, lmax, nb)
    s = capi.DeviceSeries(J)
    idx = np.array([0, 3], dtype=np.uint32)
    pidx = ctypes.c_void_p(idx.ctypes.data)
    pr = np.ascontiguousarray(starts, dtype=np.uint64)
    nn = np.ascontiguousarray(lens, dtype=np.uint32)
    narr = 14
    out_m = hip.malloc(nb * lmax * narr * 8)
    out_f = hip.malloc(nb * lmax * narr * 8)
    mean_m, var_m = hip.malloc(lmax * narr * 8), hip.malloc(lmax * narr * 8)
    mean_f, var_f = hip.malloc(lmax * narr * 8), hip.malloc(lmax * narr * 8)
    stream = hip.stream()

    def enqueue(form, out, mean, var):
        check(L.anl_gk_calculate_blocks_async(s.handle, pidx, 2, 1, lmax, 1, 0, 0, ctypes.c_void_p(pr.ctypes.data),
                                              ctypes.c_void_p(nn.ctypes.data), nb, form, out, mean, var, stream))

    tune = None
    if not args.no_autotune:
        ms_t, best = s.gk_autotune([0, 3], 1, lmax, 1, starts, lens)
        tune = {"candidates": [[8, 16, 4], [8, 16, 2], [4, 16, 4], [4, 16, 2], [8, 8, 4], [4, 8, 4]],
                "candidates_ms": [round(float(x), 3) for x in ms_t], "chosen_shape": list(s.gk_shape()),
                "shape": "tiles per wavefront, MFMA steps per FP32 chunk, wavefronts per workgroup"}
        log("gk autotune: %s -> %s" % (tune["candidates_ms"], tune["chosen_shape"]))
    enqueue(capi.GK_MFMA, out_m, mean_m, var_m)     # sizes the series scratch outside the timed region
    hip.sync()
    e0, e1, e2 = hip.event(), hip.event(), hip.event()
    reps = max(1, args.steps)
    hip.record(e0, stream)
    for _ in range(reps):
        enqueue(capi.GK_MFMA, out_m, mean_m, var_m)
    hip.record(e1, stream)
    ms_m = hip.elapsed_ms(e0, e1) / reps
    hip.record(e1, stream)
    enqueue(capi.GK_FP64, out_f, mean_f, var_f)
    hip.record(e2, stream)
    ms_f = hip.elapsed_ms(e1, e2)
    errC, errI, errK = 0.0, 0.0, 0.0
    A = np.zeros((nb, lmax, narr))
    Bf = np.zeros((nb, lmax, narr))
    hip.d2h(A, out_m, A.nbytes)
    hip.d2h(Bf, out_f, Bf.nbytes)
    for b in range(nb):
        a, b_ = A[b], Bf[b]
        c0 = np.sqrt(np.abs(np.diag(b_[0, :4].reshape(2, 2))))
        sig = np.outer(c0, c0).reshape(-1)
        errC = max(errC, float(np.max(np.abs(a[:, :4] - b_[:, :4]) / sig)))
        li = np.arange(lmax)[:, None] + 1.0
        errI = max(errI, float(np.max(np.abs(a[:, 4:12:2] - b_[:, 4:12:2]) / sig / li)))
        # k_gk = I00 - I01 I10 / I11 has poles where the sample's running
        # integral I11 crosses zero at long lags: compare where |I11| is at
        # least a tenth of its maximum
        i11 = np.abs(b_[1:, 10])
        ok = i11 >= 0.1 * i11.max()
        k = np.abs(b_[1:, 12][ok])
        errK = max(errK, float(np.max(np.abs(a[1:, 12][ok] - b_[1:, 12][ok]) / k.max())))
    mm, mf = np.zeros((lmax, narr)), np.zeros((lmax, narr))
    hip.d2h(mm, mean_m, mm.nbytes)
    hip.d2h(mf, mean_f, mf.nbytes)
    c0 = np.sqrt(np.abs(np.diag(mf[0, :4].reshape(2, 2))))
    err_mean = float(np.max(np.abs(mm[:, :4] - mf[:, :4]) / np.outer(c0, c0).reshape(-1)))
    for p in (out_m, out_f, mean_m, var_m, mean_f, var_f):
        hip.free(p)
    flop = 2.0 * 3 * 4 * float(lens.sum()) * lmax
    res = {"metric": "Green-Kubo lag contraction FLOP/s (algorithmic 2*3*ncur^2*sum(n0)*lags)",
           "value": flop / (ms_m * 1e-3), "unit": "FLOP/s", "n_gpus": 1, "steps": reps, "ms_per_step": ms_m,
           "dtype": "split-bf16 mfma, f32 chunk, f64 reduce",
           "config": {"workload": "GK 2 vector currents (6 FP64 columns), %d AR(2) steps, coupling 0.3, %d uneven "
                                  "blocks, %d lags, skip 1, n_seg 1" % (T, nb, lmax),
                      "block_lengths_min_max": [int(lens.min()), int(lens.max())], "algorithmic_flop": flop},
           "ms_all_blocks_fp64": ms_f, "speedup_mfma_vs_fp64": ms_f / ms_m, "autotune": tune,
           "roofline": {"bound": "mfma", "achieved": 3 * flop / (ms_m * 1e-3) / 1e12, "peak": BF16_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": 3 * flop / (ms_m * 1e-3) / 1e12 / BF16_PEAK_TFLOPS,
                        "note": "3 bf16 MFMA products per algorithmic product; event-timed over one call for all "
                                "32 blocks (splits, tiles, reduction, FP64 epilogue, block mean/variance)"},
           "error_vs_fp64": {"max_dC_over_sigma": errC, "max_dI_over_sigma_l": errI,
                             "max_dk_gk_over_max_k": errK, "block_mean_max_dC_over_sigma": err_mean,
                             "note": "per block; sigma_ab = sqrt(C_aa(0) C_bb(0)) of the FP64 form (itself identical "
                                     "to the oracle); k_gk error relative to the block's max |k_gk| over the lags where |I11| >= "
                                     "0.1 max |I11| (elsewhere k_gk has the sample's poles)"}}
    if not args.no_cpu:
        from oracle import oracle as O
        n_s = args.gk_cpu_n
        t = time.perf_counter()
        O.gk(J, [0, 3], 1, n_s, False, 0, 1, n_s, 0)
        dt = time.perf_counter() - t
        res["cpu_baseline"] = {"value": 2.0 * 3 * 4 * n_s * n_s / dt, "unit": "FLOP/s", "cores": 1, "kind": "port",
                               "sample": "GreenKubo reset(%d), %d lags, on the same series, %.1f s"
                                         % (n_s, n_s, dt)}
    log("gk: mfma %.2f ms, fp64 %.2f ms, max dC/sigma %.2e" % (ms_m, ms_f, errC))
    s.close()
    return res


SH_WORKLOAD = {"N": 2048, "F": 2200, "rho": 0.1, "sigma": 0.05, "seed": 71, "fractions": [0.5, 0.5],
               "rminmax": [(0.7, 3.5)] * 4, "nbin": 4, "tmax": 200, "skip": 10, "n": 2000, "lmax": 10}


def run_sh(args, hip):
    """SphericalCorrelations: per-atom spherical-harmonic densities of every
    frame, then their lag products.  Algorithmic FLOP = 2 * (number of t0) *
    lags * N * ntypes * nbin * (lmax+1)^2 (one multiply-add per harmonic
    component of every atom for every (t0, lag) pair), priced against the FP64
    peak."""
    from analisi_amd import capi, synth
    w = SH_WORKLOAD
    N, F, lmax, nbin = w["N"], w["F"], w["lmax"], w["nbin"]
    Lb = (N / w["rho"]) ** (1.0 / 3.0)
    box_l = synth.lammps_box(Lb)
    log("generating %d atoms x %d frames" % (N, F))
    pos = synth.random_walk(w["seed"], N, F, box_l, w["sigma"])
    types = synth.types_by_fraction(N, w["fractions"])
    tid, nt = type_ids(types)
    traj = capi.DeviceTrajectory(pos, internal_box(box_l, F), tid, nt)
    s = capi.DeviceSph(traj, w["rminmax"], nbin, lmax)
    n, tmax, skip = w["n"], w["tmax"], w["skip"]
    leff = min(n, tmax) if tmax else n
    n0 = (n + skip - 1) // skip
    out = s.calculate(n, tmax, skip, 0)             # untimed: workspace, code objects
    hip.sync()
    e0, e1 = hip.event(), hip.event()
    reps = max(1, args.steps)
    hip.record(e0, None)
    for _ in range(reps):
        out = s.calculate(n, tmax, skip, 0)
    hip.record(e1, None)
    ms = hip.elapsed_ms(e0, e1) / reps
    flop = 2.0 * n0 * leff * N * nt * nbin * (lmax + 1) ** 2
    # lag-0 check: every l=0 entry is mean_i (count_i / sqrt(4 pi))^2 > 0
    assert np.all(out[0, ..., lmax] > 0)
    res = {"metric": "spherical-harmonic density lag products FLOP/s (algorithmic 2*n0*lags*N*ntypes*nbin*(lmax+1)^2)",
           "value": flop / (ms * 1e-3), "unit": "FLOP/s", "n_gpus": 1, "steps": reps, "ms_per_step": ms,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "SphericalCorrelations lmax %d: %d atoms 50/50, cubic %.2f A (0.1 atoms/A^3), %d "
                                  "frames, sigma %g A (seed %d); ranges (0.7, 3.5) A, %d bins, %d lags, t0 stride %d, "
                                  "reset(%d), calculate(0)" % (lmax, N, Lb, F, w["sigma"], w["seed"], nbin, leff, skip,
                                                               n),
                      "algorithmic_flop": flop},
           "roofline": {"bound": "fp64", "achieved": flop / (ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                        "unit": "TFLOP/s", "frac": flop / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                        "note": "event-timed over whole calculate() calls (densities of every frame, lag-product "
                                "tiles, t0 reduction, copy of the result)"}}
    if not args.no_cpu:
        from oracle import oracle as O
        ot = O.Traj(pos, types, np.tile(box_l, (F, 1)), False)
        nf = 2
        t = time.perf_counter()
        dens = np.stack([O.sph_density(ot, f, w["rminmax"], nbin, lmax) for f in range(nf + 1)])
        t_frame = (time.perf_counter() - t) / (nf + 1)
        t = time.perf_counter()
        O.sph_corr_from_densities(ot, dens, nbin, 2, 1, 2, lmax)
        t_pair = (time.perf_counter() - t) / 4
        nframes = n + leff - 1
        est = nframes * t_frame + n0 * leff * t_pair
        res["cpu_baseline"] = {"value": flop / est, "unit": "FLOP/s", "cores": 1, "kind": "port",
                               "sample": "oracle on the same trajectory: density of one frame %.3f s (mean of %d), one "
                                         "(t0, lag) product %.4f s (mean of 4); the whole workload's %d frames and "
                                         "%d (t0, lag) pairs extrapolate to %.0f s" %
                                         (t_frame, nf + 1, t_pair, nframes, n0 * leff, est)}
    # Steinhardt histogram of (q4, q6), 100 bins each, over every frame
    ll, nbs = [4, 6], 100
    s.steinhardt(F, 1, 0, ll, nbs)                  # untimed
    hip.sync()
    hip.record(e0, None)
    for _ in range(reps):
        hst = s.steinhardt(F, 1, 0, ll, nbs)
    hip.record(e1, None)
    ms_st = hip.elapsed_ms(e0, e1) / reps
    assert hst.sum() > 0
    res["steinhardt"] = {"workload": "SteinhardtOrderParameterHistogram_10 of the same trajectory: (q4, q6) "
                                     "histogram, 100 x 100 cells, every one of the %d frames, 4 bins" % F,
                         "ms_per_step": ms_st, "frames_per_s": F / (ms_st * 1e-3)}
    if not args.no_cpu:
        t = time.perf_counter()
        O.steinhardt(ot, w["rminmax"], nbin, lmax, nbs, ll, 1, 2, 0, True)
        t_st = (time.perf_counter() - t) / 2
        res["steinhardt"]["cpu_baseline"] = {"value": 1.0 / t_st, "unit": "frames/s", "cores": 1, "kind": "port",
                                             "sample": "oracle, 2 frames of the same trajectory, %.3f s per frame"
                                                       % t_st}
    log("sh: %.2f ms per calculate, steinhardt %.2f ms for %d frames" % (ms, ms_st, F))
    s.close()
    traj.close()
    return res


def run_lammps_feed(args, hip):
    """gofrt_block_average over a memory-mapped LAMMPS-binary trajectory
    (Traj): each block moves the window; with read-ahead the next block's
    window is read, ordered and uploaded on a host thread while the current
    block's kernels run.  Both modes give bit-identical means and variances."""
    import tempfile
    from analisi_amd import pyanalisi as pa, synth
    N, F, L, nb, tmax = 8192, 600, 45.0, 6, 41
    box_l = synth.lammps_box(L)
    pos = synth.random_walk(5, N, F, box_l, 0.05)
    types = synth.types_by_fraction(N, [0.8, 0.1, 0.1]).astype(np.int32)
    t = pa.Trajectory(pos, np.zeros_like(pos), types, np.tile(box_l, (F, 1)), pa.BoxFormat.LammpsOrtho, False, False)
    fd, path = tempfile.mkstemp(suffix=".bin")
    os.close(fd)
    try:
        t.write_lammps_binary(path, 0, -1)
        del t
        tl = pa.Traj(path)
        s = (F - min(F // (nb + 1) + 1, tmax)) // nb
        tl.setAccessWindowSize(s + tmax - 1)
        tl.setAccessStart(0)
        times, outs = {True: [], False: []}, {}
        for mode in [True, False] * max(1, args.steps):
            tl.setPrefetch(mode)
            t0 = time.perf_counter()
            outs[mode] = pa.gofrt_block_average(tl, nb, 0.0, 12.0, 256, tmax, 1, 1)
            times[mode].append(time.perf_counter() - t0)
        same = all(np.array_equal(np.nan_to_num(a), np.nan_to_num(b)) for a, b in zip(outs[True], outs[False]))
        size = os.path.getsize(path)
    finally:
        os.remove(path)
    on, off = min(times[True]), min(times[False])
    log("lammps feed: read-ahead %.3f s, without %.3f s" % (on, off))
    return {"metric": "g(r,t) block average over a LAMMPS-binary file, wall time", "value": on, "unit": "s",
            "n_gpus": 1, "steps": len(times[True]), "higher_is_better": False, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "gofrt_block_average(Traj, %d blocks of %d t0, lags 0..%d, 256 bins to 12 A): "
                                   "%d atoms 80/10/10, %d frames, cubic %g A, sigma 0.05 A (seed 5), %d MiB file, "
                                   "window %d frames" % (nb, s, tmax - 1, N, F, L, size >> 20, s + tmax - 1)},
            "without_read_ahead_s": off, "speedup": off / on, "bit_identical": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5", "sh", "lammps"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune-lags", type=int, default=4, help="lags per autotune probe launch")
    ap.add_argument("--shape", default="", help="g(r,t) launch shape 'unroll,xcd,log2 pairs per workgroup,cluster' "
                                                "instead of the autotuner (for reproducible profiles)")
    ap.add_argument("--cpu-n", type=int, default=0,
                    help="reset(n) of the CPU baseline sample (n t0 x n lags; 0 = 2 on c3, 8 on c2, 16 on a c4 window)")
    ap.add_argument("--c4-windows", type=int, default=0, help="number of c4 windows (0 = the workload's 64)")
    ap.add_argument("--c4-split", type=int, default=1, help="c4 windows cut into this many graphs per step")
    ap.add_argument("--gk-cpu-n", type=int, default=6000, help="reset(n), n lags, of the GK oracle sample")
    ap.add_argument("--msd-cpu-n", type=int, default=20, help="reset(n) of the MSD oracle sample")
    ap.add_argument("--fp64-n", type=int, default=0, help="reset(n) of the plain-FP64 slice (0 = 101 on c2, 16 else)")
    ap.add_argument("--no-gk", action="store_true", help="c3: leave out the Green-Kubo (c5) sub-record")
    ap.add_argument("--n", type=int, default=0, help="reset(n) of the timed run (0 = the workload's own)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain streaming kernel that reads and writes MIB mebibytes (to scale the "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        # a plain streaming kernel of known traffic: anl_counts_to_double reads
        # one uint64 and writes one double per element (MIB MiB each way)
        from analisi_amd import capi
        hip = Hip()
        capi.check(capi.lib().anl_init(0))
        nb = args.calibrate_copy << 20
        src, dst = hip.malloc(nb), hip.malloc(nb)
        for _ in range(3):
            capi.check(capi.lib().anl_counts_to_double_async(src, 1.0, nb // 8, dst, None))
        hip.sync()
        print(json.dumps({"kernel": "counts_to_double_kernel", "read_bytes": nb, "write_bytes": nb, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: analisi_amd has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    from analisi_amd import capi
    hip = Hip()
    capi.check(capi.lib().anl_init(0))
    if args.config == "c5":
        res = run_gk(args, hip)
    elif args.config == "sh":
        res = run_sh(args, hip)
    elif args.config == "lammps":
        res = run_lammps_feed(args, hip)
    else:
        w = gofrt_workload(args.config)
        if args.n:
            w["n"] = args.n
        res = run_blocks(w, args, hip) if args.config == "c4" else run_gofrt(w, args, hip)
        if args.config == "c3" and not args.no_gk:
            res["gk"] = run_gk(args, hip)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
