"""Benchmark of the per-step PM loop on one MI355X.

Workload (default): 1024^3 particles on a 1024^3 grid, periodic box 2000 Mpc/h,
Zel'dovich initial conditions at z = 50 (Eisenstein-Hu no-wiggle P(k), sigma_8 = 0.81,
Philox white noise, seed 2), 50 a-uniform steps to z = 0; FP32 engine (int32 cell +
FP32 offset positions, int64 fixed-point deposit, complex64 FFT).  A step is one pass of
drift -> CIC deposit -> FFT Poisson/gradient solve -> gather/kick -> drift over every
particle, with inputs resident in HBM.  W warmup steps, then K timed steps bracketed by
device synchronisation; the value is particle-steps per second.

Also reported on the same line:
  roofline     : the dominant kernel (largest share of the timed region, timed with HIP
                 events on the engine stream) against MI355X HBM bandwidth, from its
                 algorithmic bytes per launch;
  accuracy     : the whole run on the FP32 engine and on the plain FP64 GPU path
                 (reference-structured kernels in FP64) from the same initial
                 conditions: max relative P(k) error for k < k_Nyq/2, RMS displacement
                 (grid units) and velocity errors, the density bit-exact across deposits;
  fp64_plain   : the FP64 path's rate over that run, and the FP32/FP64 speed-up;
  cpu_baseline : the reference algorithm restated on the host CPU (the oracle) on the
                 64^3-particle, 120-step workload, in the reference's FP32 rounding and
                 in FP64, on all host threads and on one, per-stage seconds;
  ensemble     : 96 boxes (32 each of 64^3, 128^3 and 192^3 particles, per-box seed and
                 sigma_8) stepped 40 steps z = 30 -> 0 as a batched ensemble (HIP graph
                 runs) and one box at a time, and the speed-up;
  pressure_ensemble : 24 boxes of 384^3 particles on 768^3 grids with 200 of 96^3,
                 30 steps z = 40 -> 0, in 90% of the free device memory (the large
                 batch in waves).
"""
import argparse
import json
import os
import sys
import subprocess
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s, 1024^3 PM FP32 vs FP64 at stated P(k) err; % HBM roofline"
HBM_PEAK = 8.0e12


def kernel_bytes(name, Np, G):
    """Algorithmic bytes per launch of each timed kernel region: what the kernel must
    read and write once (no re-reads, no atomics beyond one write per cell)."""
    table = {
        "drift": 36 * Np,                      # read pos 12 + vel 12, write pos 12
        "zero_density": 8 * G,                 # write int64 density
        "deposit": 12 * Np + 4 * G,            # read pos (12 B fixed point), write the FP32 density of the step
        "deposit_windows": 12 * Np,            # read pos (the window arena is overhead)
        "deposit_owner": 4 * G,                # write the FP32 density (the step path; int64 for hgpm_cic)
        # half spectra (k_z = 0..N/2) cost 4 B per grid cell
        "fft_fwd_z_r2c": 8 * G,                # read the FP32 density, write the half spectrum
        "fft_fwd_y": 8 * G,                    # read + write one half spectrum
        "fft_x_psi": 8 * G,                    # read rho^, write psi^
        "fft_inv_y": 8 * G,                    # read + write one half spectrum
        "fft_inv_z_c2r": 8 * G,                # read psi^, write psi (FP32)
        "gather_kick_drift": 48 * Np + 4 * G,  # pos (12 B) + vel (12 B) read and write, potential read
        "gather_kick": 36 * Np + 4 * G,        # pos + vel read, vel write, potential read
    }
    return table.get(name)


def measured_traffic(workload, region):
    """HBM bytes per launch of a timed region, from the committed counter summaries
    (profiles/*.md lines `traffic: <workload>/steps<first>-<last> <region> <bytes>`, the
    counter passes over the same timed steps; FETCH_SIZE and WRITE_SIZE passes scaled by
    a plain copy of known size).  None if not measured."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for fn in sorted(os.listdir(d)):
        if not fn.endswith(".md"):
            continue
        for line in open(os.path.join(d, fn)):
            f = line.split()
            if len(f) == 4 and f[0] == "traffic:" and f[1] == workload and f[2] == region:
                return float(f[3])
    return None


def make_ics(hg, p, ts):
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "eh.ipk")
        hg.write_ipk_eh(p, fn, 100.0, 200001)
        hdr, vals = hg.read_ipk(fn)
    return hdr, vals


def progress(msg):
    """A progress line on stderr (stdout carries the one JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_engine(hg, p, steps, warmup, timing=True, gather_tile=-1):
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    hdr, vals = make_ics(hg, p, ts)
    n_total = warmup + steps
    sched = hg.schedule(p, n_total, merge=True)
    with hg.MemoryManager(p) as m:
        t0 = time.perf_counter()
        hg.GenerateDisplacementIC(m, p, ts, hdr, vals)
        m.synchronize()
        t_ic = time.perf_counter() - t0
        # FFT shapes, then the deposit tile and gather shapes on these particles
        tune = m.autotune(3) if p.precision == hg.FP32 else None
        if gather_tile >= 0 and p.precision == hg.FP32:
            m.set_tunable("gather_tile", gather_tile)
        for s in range(warmup):
            m.step(sched[s])
        m.synchronize()
        if timing:
            m.set_timing(True)
            m.reset_timings()
        m.synchronize()
        t0 = time.perf_counter()
        for s in range(warmup, n_total):
            m.step(sched[s])
        m.synchronize()
        el = time.perf_counter() - t0
        tim = m.timings() if timing else {}
        dev_bytes = m.device_bytes()
    return el, tim, t_ic, dev_bytes, tune


def run_full(hg, p, per_kernel=True):
    """The whole run of the workload, z_ini -> z_fin in p.nsteps steps from fresh initial
    conditions, each step timed on its own (host clock, the stream synchronised after
    every step): clustering grows toward z = 0, and with it the deposit's windows.
    per_kernel: the kernel regions' HIP-event times are averaged over the first and the
    last 10% of the steps as well."""
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    hdr, vals = make_ics(hg, p, ts)
    n = p.nsteps
    sched = hg.schedule(p, n, merge=True)
    t = []
    deferred = []
    arena = []  # 32-bit arena words reserved per grid cell, per step
    regions = []
    with hg.MemoryManager(p) as m:
        hg.GenerateDisplacementIC(m, p, ts, hdr, vals)
        m.autotune(3)
        m.synchronize()
        for s in range(n):
            m.set_timing(per_kernel)
            m.reset_timings()
            t0 = time.perf_counter()
            m.step(sched[s])
            m.synchronize()
            t.append(time.perf_counter() - t0)
            regions.append(m.timings() if per_kernel else {})
            deferred.append([m.get_tunable(k) for k in
                             ("deposit_deferred", "deposit_wide", "deposit_deferred_full")])
            arena.append(m.get_tunable("deposit_arena_mwords") / (p.ng ** 3 / 2 ** 20))
    k = max(1, n // 10)
    Np = p.np ** 3

    def region_ms(rs):
        names = sorted({name for r in rs for name in r})
        return {name: round(sum(r.get(name, (0.0, 0))[0] for r in rs) / len(rs), 3) for name in names}

    return {"steps": n, "z": [p.z_ini, p.z_fin], "ms_per_step_mean": 1e3 * sum(t) / n,
            "ms_per_step_first_10pct": 1e3 * sum(t[:k]) / k, "ms_per_step_last_10pct": 1e3 * sum(t[-k:]) / k,
            "particle_steps_per_s": Np * n / sum(t), "deferred_tiles_max": max(d[0] for d in deferred),
            # per step: tiles deferred, tiles with 64-bit windows, tiles deferred for a full
            # slot list or sub-arena
            "deferred_tiles": deferred,
            # every window reservation (those of tiles then found full and the doubled ones
            # of the 64-bit pass included), in 32-bit words per grid cell
            "arena_words_per_grid_cell": [round(w, 3) for w in arena],
            "per_kernel": per_kernel,
            "kernel_ms_first_10pct": region_ms(regions[:k]), "kernel_ms_last_10pct": region_ms(regions[-k:]),
            "ms_per_step": [round(1e3 * x, 3) for x in t]}


def copy_ceiling(gib=8, reps=3):
    """Streaming rate a plain elementwise kernel reaches on this device (torch
    out = a + 0 over gib GiB, read + write): the practical ceiling beside the published
    8 TB/s."""
    import torch
    n = (gib << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = 0.0
    for _ in range(reps):
        e0.record()
        torch.add(a, 0.0, out=b)
        e1.record()
        torch.cuda.synchronize()
        best = max(best, 2.0 * 4 * n / (e0.elapsed_time(e1) * 1e-3) / 1e12)
    del a, b
    torch.cuda.empty_cache()
    return {"copy_TBps": best, "bytes_read": 4 * n, "bytes_written": 4 * n,
            "note": "torch add(a, 0, out=b), best of %d" % reps}


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", read from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(hg, threads):
    """The reference algorithm on the host: the oracle's restatement of the per-step PM
    loop on the CPU-runnable workload, 64^3 particles on a 64^3 grid, box 128 Mpc/h,
    Zel'dovich initial conditions at z = 200 (seed 21082001, the reference's own
    test_run_params), 120 a-uniform steps to z = 0 (with 60 steps the reference's
    initial half step back makes a < 0).  Two precisions: the reference's FP32 rounding
    (float3 positions, FP32 CIC sums in a fixed order, the reference's calcGreens
    rounding; `value`) and FP64.  On `threads` OpenMP threads, and once on 1 thread."""
    from oracle import oracle as orc
    ng, n = 64, 120
    p = hg.make_params(ng, ng, 128.0, 200.0, 0.0, n, seed=21082001)
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    hdr, vals = make_ics(hg, p, ts)
    a = 1 / (1 + p.z_ini)
    D, _ = hg.growth(p, a)
    _, dD = hg.growth(p, a - ts.deltaT / 2)
    pos, vel = orc.ic_f64(ng, ng, p.rl, p.seed, vals, hdr[3], D, dD, p.z_ini, ts.fscal, ts.deltaT)
    co = hg.schedule_array(p, n, merge=False)
    Np = ng ** 3
    runs = {}
    for label, fn, th in [("ref32", orc.run_ref32, threads), ("f64", orc.run_f64, threads),
                          ("ref32_1thread", orc.run_ref32, 1)]:
        orc.set_threads(th)
        st = {}
        t0 = time.perf_counter()
        if label.startswith("ref32"):
            fn(ng, ng, co, pos.astype(np.float32), vel.astype(np.float32), st)
        else:
            fn(ng, ng, co, pos, vel, st)
        el = time.perf_counter() - t0
        runs[label] = {"threads": th, "seconds": el, "particle_steps_per_s": Np * n / el,
                       "stage_seconds": {k: round(v, 3) for k, v in st.items()}}
    orc.set_threads(threads)
    return {"value": runs["ref32"]["particle_steps_per_s"], "unit": "particle-steps/s", "cores": threads,
            "kind": "port", "model": cpu_model(), "threads": threads,
            "value_1thread": runs["ref32_1thread"]["particle_steps_per_s"],
            "value_f64": runs["f64"]["particle_steps_per_s"],
            "per_stage": runs["ref32"]["stage_seconds"], "runs": runs,
            "sample": f"64^3 particles, 64^3 grid, box 128 Mpc/h, z=200->0 in {n} steps, seed 21082001: the "
                      f"oracle in the reference's FP32 rounding (value) and in FP64, {threads} threads, "
                      f"and the FP32 rounding on 1 thread"}


def ensemble_boxes(hg, per_class=32, sizes=(64, 128, 192), z_ini=30.0, n_steps=40):
    """The batched-ensemble workload: per_class boxes at each size (grid = particles per
    dimension, 2 Mpc/h per cell), seed 1000+b, sigma_8 drawn uniformly in [0.7, 0.9]
    by a generator seeded 4242, z_ini -> 0 in n_steps."""
    rng = np.random.default_rng(4242)
    ps = []
    for n in sizes:
        for _ in range(per_class):
            b = len(ps)
            ps.append(hg.make_params(n, n, 2.0 * n, z_ini, 0.0, n_steps, seed=1000 + b,
                                     ss8=float(rng.uniform(0.7, 0.9))))
    return ps


def pressure_boxes(hg, n_big=24, n_small=200, z_ini=40.0, n_steps=30):
    """The memory-pressure workload: n_big boxes of 384^3 particles on 768^3 grids (box
    1500 Mpc/h) with n_small boxes of 96^3 particles on 96^3 grids (box 192 Mpc/h),
    seeds 2000 + b, z_ini -> 0 in n_steps."""
    ps = [hg.make_params(768, 384, 1500.0, z_ini, 0.0, n_steps, seed=2000 + b) for b in range(n_big)]
    ps += [hg.make_params(96, 96, 192.0, z_ini, 0.0, n_steps, seed=2000 + n_big + b) for b in range(n_small)]
    return ps


def run_pressure(hg, steps, warmup, n_big, n_small, budget, classes_alone=True):
    """The memory-pressure ensemble in one device budget (0: 90% of the free memory):
    batches that do not fit keep grid buffers for a wave of their boxes and step the
    waves in turn.  Then each size class alone in the same budget."""
    ps = pressure_boxes(hg, n_big, n_small)
    key = lambda p: (p.ng, p.np)
    ics, scheds = {}, {}
    for p in ps:
        if key(p) not in ics:  # the classes share their cosmology; the seeds are per box
            ics[key(p)] = ic_args(hg, p)
            scheds[key(p)] = hg.schedule(p, warmup + steps, merge=True)

    def one(pl):
        with hg.Ensemble(pl, device_budget=budget) as e:
            t0 = time.perf_counter()
            for b, p in enumerate(pl):
                e.generate_ic(b, *ics[key(p)])
            e.synchronize()
            t_ic = time.perf_counter() - t0
            e.autotune(2)
            for s in range(warmup):
                e.step([scheds[key(p)][s] for p in pl])
            e.synchronize()
            t0 = time.perf_counter()
            for s in range(warmup, warmup + steps):
                e.step([scheds[key(p)][s] for p in pl])
            e.synchronize()
            el = time.perf_counter() - t0
            nb, of = e.batches()
            waves = [int(w) for w in e.waves()]
            sizes = [int(sum(1 for x in of if x == g)) for g in range(nb)]
            return {"ms_per_step": el * 1e3 / steps,
                    "particle_steps_per_s": sum(p.np ** 3 for p in pl) * steps / el,
                    "batches": [{"np": pl[list(of).index(g)].np, "ng": pl[list(of).index(g)].ng,
                                 "boxes": sizes[g], "boxes_per_wave": waves[g]} for g in range(nb)],
                    "device_GB": e.device_bytes() / 1e9, "ic_seconds": t_ic}

    progress(f"{len(ps)} boxes in one ensemble")
    mixed = one(ps)
    alone = []
    if classes_alone:
        progress("each size class alone")
        alone = [dict(one([p for p in ps if key(p) == k]), np=k[1], ng=k[0]) for k in sorted(ics)]
    nparts = sum(p.np ** 3 for p in ps)
    return {"metric": "aggregate particle-steps/s, ensemble under memory pressure",
            "value": mixed["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": mixed["ms_per_step"], "higher_is_better": True,
            "mixed": mixed, "per_class_alone": alone, "dtype": "f32",
            "data": "synthetic: Zel'dovich ICs at z=40, Eisenstein-Hu no-wiggle P(k), per-box seed",
            "config": {"workload": f"pressure_{n_big}x(384^3_on_768^3_box1500)+{n_small}x(96^3_box192)_z40",
                       "boxes": len(ps), "particles": nparts, "device_budget_GB": budget / 1e9}}


def ic_args(hg, p):
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "eh.ipk")
        hg.write_ipk_eh(p, fn, 20.0, 20001)
        hdr, vals = hg.read_ipk(fn)
    a = 1.0 / (1.0 + p.z_ini)
    delta, _ = hg.growth(p, a)
    _, dot_delta = hg.growth(p, a - ts.deltaT / 2.0)
    return hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT


def run_ensemble(hg, steps, warmup, per_class, classes_alone=True):
    """Batched ensemble, its steps captured into one HIP graph, vs the same boxes
    launched one at a time (one context per box, each box's step completed before the
    next box's starts).  The batched run with eager launches is reported beside it."""
    ps = ensemble_boxes(hg, per_class)
    n_total = warmup + steps
    ics = [ic_args(hg, p) for p in ps]
    scheds = [hg.schedule(p, n_total, merge=True) for p in ps]
    nparts = sum(p.np ** 3 for p in ps)
    with hg.Ensemble(ps) as e:
        for b, a in enumerate(ics):
            e.generate_ic(b, *a)
        e.autotune(2)
        for s in range(warmup):
            e.step([sc[s] for sc in scheds])
        e.synchronize()
        t0 = time.perf_counter()
        for s in range(warmup, n_total):
            e.step([sc[s] for sc in scheds])
        e.synchronize()
        t_eager = time.perf_counter() - t0
        nbatches, _ = e.batches()
        dev = e.device_bytes()
    # the same run with the steps captured into one HIP graph (capture and instantiation
    # inside the timed region)
    with hg.Ensemble(ps) as e:
        for b, a in enumerate(ics):
            e.generate_ic(b, *a)
        e.autotune(2)
        e.run([[sc[s] for sc in scheds] for s in range(warmup)])
        e.synchronize()
        t0 = time.perf_counter()
        e.run([[sc[s] for sc in scheds] for s in range(warmup, n_total)])
        e.synchronize()
        t_batched = time.perf_counter() - t0
    progress("one box at a time")
    t_single = 0.0
    tuned = {}  # per size class: the autotuner's choices on its first box, kept for the others
    keys = ("fft_plan", "fft_cols", "fft_xsolve", "deposit_tile", "gather_tile")
    for b, p in enumerate(ps):
        with hg.MemoryManager(p) as m:
            m.generate_ic(*ics[b])
            if p.np not in tuned:
                m.autotune(2)
                tuned[p.np] = {k: m.get_tunable(k) for k in keys}
            else:
                for k, v in tuned[p.np].items():
                    m.set_tunable(k, v)
            for s in range(warmup):
                m.step(scheds[b][s])
            m.synchronize()
            t0 = time.perf_counter()
            for s in range(warmup, n_total):
                m.step(scheds[b][s])
                m.synchronize()
            t_single += time.perf_counter() - t0
    # each size class batched on its own (no other class sharing the GPU)
    classes = []
    if classes_alone:
        progress("each size class alone")
    for n in sorted({p.np for p in ps}) if classes_alone else []:
        idx = [b for b, p in enumerate(ps) if p.np == n]
        with hg.Ensemble([ps[b] for b in idx]) as e:
            for j, b in enumerate(idx):
                e.generate_ic(j, *ics[b])
            e.autotune(2)
            for s in range(warmup):
                e.step([scheds[b][s] for b in idx])
            e.synchronize()
            t0 = time.perf_counter()
            for s in range(warmup, n_total):
                e.step([scheds[b][s] for b in idx])
            e.synchronize()
            t = time.perf_counter() - t0
        classes.append({"np": n, "boxes": len(idx), "ms_per_step": t * 1e3 / steps,
                          "particle_steps_per_s": len(idx) * n ** 3 * steps / t})
    return {"metric": "aggregate particle-steps/s, batched ensemble vs one box at a time",
            "value": nparts * steps / t_batched, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": t_batched * 1e3 / steps, "higher_is_better": True,
            "one_at_a_time": {"value": nparts * steps / t_single, "ms_per_step": t_single * 1e3 / steps},
            "batched_eager": {"value": nparts * steps / t_eager, "ms_per_step": t_eager * 1e3 / steps},
            "speedup": t_single / t_batched, "batches": nbatches, "device_GB": dev / 1e9,
            "per_class_batched_alone": classes,
            "dtype": "f32", "data": "synthetic: Zel'dovich ICs, Eisenstein-Hu no-wiggle P(k), per-box seed and sigma_8",
            "config": {"workload": f"ensemble_{per_class}x(64^3,128^3,192^3)_box2Mpc_per_cell_z30",
                       "boxes": len(ps), "particles": nparts}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the single-box run is 50 steps from z = 50 to z = 0 and the ensemble's 40 steps from
    # z = 30 to z = 0: by default the 3 first are the warm-up and the others (through
    # z = 0) are timed
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ng", type=int, default=1024)
    ap.add_argument("--rl", type=float, default=2000.0)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble comparison")
    ap.add_argument("--no-pressure", action="store_true", help="skip the memory-pressure ensemble")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the whole z_ini -> z_fin run timed step by step")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB device-to-device twice (a plain copy kernel of known "
                         "size, the reference point for the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--gather-tile", type=int, default=-1, help="gather workgroup shape (library default: -1)")
    ap.add_argument("--ensemble", action="store_true",
                    help="run the batched-ensemble workload instead (its own JSON line)")
    ap.add_argument("--per-class", type=int, default=32, help="ensemble boxes per size class")
    ap.add_argument("--pressure", action="store_true",
                    help="run the memory-pressure ensemble (24 x 384^3 on 768^3 with 200 x 96^3) instead")
    ap.add_argument("--pressure-boxes", type=int, nargs=2, default=[24, 200], metavar=("BIG", "SMALL"))
    ap.add_argument("--budget-gb", type=float, default=0.0, help="device budget of --pressure (0: 90%% free)")
    ap.add_argument("--only-ceiling", action="store_true",
                    help="print the streaming copy rate of a plain elementwise kernel and exit")
    ap.add_argument("--only-copy", action="store_true",
                    help="run only the --copy-calibration copy (a separate profiler pass)")
    args = ap.parse_args()
    if args.copy_calibration and not args.only_copy:
        # torch's HIP runtime and libhgpm's must not share a process: the calibration
        # copy runs alone, as its own profiler pass
        ap.error("--copy-calibration runs only with --only-copy (a process of its own)")
    if args.only_ceiling:
        print(json.dumps(copy_ceiling()))
        return
    if args.only_copy:
        import torch
        n = max(args.copy_calibration, 1) << 20
        src = torch.ones(n // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs a single GPU (one rank); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    import hacc_gpm_amd as hg

    if args.pressure:
        steps = args.steps if args.steps is not None else max(1, 30 - args.warmup)
        print(json.dumps(run_pressure(hg, steps, args.warmup, args.pressure_boxes[0], args.pressure_boxes[1],
                                      int(args.budget_gb * 1e9))))
        return
    if args.ensemble:
        steps = args.steps if args.steps is not None else max(1, 40 - args.warmup)
        print(json.dumps(run_ensemble(hg, steps, args.warmup, args.per_class)))
        return
    if args.steps is None:
        args.steps = max(1, 50 - args.warmup)
    ng = args.ng
    workload = f"pm_step_{ng}^3_particles_{ng}^3_grid_box{int(args.rl)}"
    n_run = max(50, args.warmup + args.steps) # steps of the whole run, z = 50 -> 0
    p = hg.make_params(ng, ng, args.rl, 50.0, 0.0, n_run, seed=2, pk_bins=221)
    Np = ng ** 3
    G = ng ** 3
    progress(f"{workload}: {args.warmup} warm-up and {args.steps} timed steps")
    el, tim, t_ic, dev_bytes, tune = run_engine(hg, p, args.steps, args.warmup, gather_tile=args.gather_tile)
    progress(f"timed steps done: {el * 1e3 / args.steps:.2f} ms per step")
    value = Np * args.steps / el
    ms_step = el * 1e3 / args.steps
    # dominant kernel region
    best = max(tim.items(), key=lambda kv: kv[1][0])
    name, (ms_tot, launches) = best
    per_launch = kernel_bytes(name, Np, G)
    avg_ms = ms_tot / max(launches, 1)
    achieved = per_launch / (avg_ms * 1e-3) if per_launch else None
    breakdown = {k: {"ms_per_step": v[0] / args.steps, "launches": v[1],
                     "GBps": (kernel_bytes(k, Np, G) / (v[0] / v[1] * 1e-3) / 1e9) if kernel_bytes(k, Np, G) and v[1] else None}
                 for k, v in sorted(tim.items(), key=lambda kv: -kv[1][0])}
    # algorithmic bytes of one step of this algorithm: 60 per particle (deposit read of
    # the 12-byte fixed-point position, gather/kick/drift read and write of position and
    # velocity 48) and 48 per grid cell (FP32 density write 4, the five half-spectrum
    # passes 8 each, potential read 4).  With 24-byte positions, an int64 density and a
    # solve that transforms the three force fields a step moves 96 Np + 120 G with half
    # spectra and 96 Np + 200 G with full complex spectra: the same step's rate against
    # those byte counts is reported beside it.
    step_alg = 60 * Np + 48 * G
    step_r2c = 96 * Np + 120 * G
    step_c2c = 96 * Np + 200 * G
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Zel'dovich ICs at z=50 from Philox white noise and an Eisenstein-Hu no-wiggle P(k)",
        "config": {"workload": workload, "np": ng, "ng": ng,
                   "rl": args.rl, "z_ini": 50, "z_fin": 0, "steps_in_run": n_run,
                   "timed_steps": [args.warmup, args.warmup + args.steps - 1], "positions": "fixed point, 3 x uint32 in 2^-22 cells",
                   "deposit": "int64 fixed point 2^-24"},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved / 1e9 if achieved else None,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK if achieved else None,
                     "traffic": measured_traffic(f"{workload}/steps{args.warmup}-{args.warmup + args.steps - 1}", name),
                     "alg_bytes_per_launch": per_launch, "avg_launch_ms": avg_ms},
        "step_roofline": {"alg_bytes_per_step": step_alg, "achieved_GBps": step_alg / (ms_step * 1e-3) / 1e9,
                          "frac": step_alg / (ms_step * 1e-3) / HBM_PEAK,
                          "force_field_solve_r2c": {"bytes_per_step": step_r2c,
                                                    "frac": step_r2c / (ms_step * 1e-3) / HBM_PEAK},
                          "force_field_solve_c2c": {"bytes_per_step": step_c2c,
                                                    "frac": step_c2c / (ms_step * 1e-3) / HBM_PEAK}},
        "kernels": breakdown,
        "autotune": tune,
        "ic_seconds": t_ic,
        "device_GB": dev_bytes / 1e9,
    }
    if not args.no_full_run:
        progress("whole run, step by step")
        out["full_run"] = run_full(hg, p)
        progress("streaming ceiling (child process)")
        # torch and this library do not share one HIP runtime in a process: a child
        # process measures the ceiling
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--only-ceiling"], capture_output=True,
                           text=True, timeout=300)
        out["hbm_streaming_ceiling"] = (json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0
                                        else {"error": r.stderr.strip()[-300:]})
    if not args.no_fp64:
        # the accuracy clause of the metric, on the same workload: the whole run on the
        # FP32 engine and on the FP64 plain path from the FP64 path's initial conditions
        # (hacc_gpm_amd.accuracy); the FP64 path's rate over that run is the plain GPU
        # baseline of the speed-up
        try:
            from hacc_gpm_amd import accuracy
            progress(f"accuracy: whole run on the FP32 engine and the FP64 plain path ({n_run} steps)")
            acc = accuracy.compare(ng, ng, args.rl, 50.0, n_run, 2, pk_bins=221, quiet=True)
            out["accuracy"] = {k: acc[k] for k in ("workload", "pk_max_rel_err", "pk_bins_compared",
                                                   "pk_counts_equal", "rms_disp_cells", "max_disp_cells",
                                                   "vel_rel_rms", "density_bit_exact", "ms_per_step_fp32",
                                                   "ms_per_step_fp64")}
            out["accuracy"]["pk_err_bound"] = 1e-3
            v64 = Np * 1e3 / acc["ms_per_step_fp64"]
            out["fp64_plain"] = {"value": v64, "ms_per_step": acc["ms_per_step_fp64"], "steps": n_run}
            out["speedup_vs_fp64"] = value / v64
        except Exception as e:  # report, do not hide
            out["accuracy"] = {"error": str(e)}
    if not args.no_ensemble:
        # the batched-ensemble workload (96 boxes of 64^3, 128^3 and 192^3, 40 steps
        # z = 30 -> 0) against the same boxes one at a time
        progress("ensemble: 3 x 32 boxes batched (HIP graph), then one at a time")
        try:
            ens = run_ensemble(hg, max(1, 40 - 3), 3, 32, classes_alone=False)
            out["ensemble"] = {k: ens[k] for k in ("value", "unit", "ms_per_step", "one_at_a_time", "batched_eager",
                                                   "speedup", "batches", "config")}
        except Exception as e:  # report, do not hide
            out["ensemble"] = {"error": str(e)}
    if not args.no_pressure:
        # the memory-pressure workload (24 x 384^3 on 768^3 grids with 200 x 96^3, 30 steps
        # z = 40 -> 0) in 90% of the free device memory: batches stepped in waves
        progress("pressure ensemble: 24 x 384^3 on 768^3 + 200 x 96^3")
        try:
            pr = run_pressure(hg, 27, 3, 24, 200, 0, classes_alone=False)
            out["pressure_ensemble"] = {k: pr[k] for k in ("value", "unit", "ms_per_step", "mixed", "config")}
        except Exception as e:  # report, do not hide
            out["pressure_ensemble"] = {"error": str(e)}
    if not args.no_cpu:
        th = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        progress(f"CPU baseline on {th} threads")
        out["cpu_baseline"] = cpu_baseline(hg, th)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
