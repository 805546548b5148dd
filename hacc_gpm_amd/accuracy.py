"""FP32 engine against the FP64 plain path, from identical initial conditions.

The FP64 path (the reference's kernels restated in double precision) generates the
Zel'dovich initial conditions; the same positions and velocities are handed to the
FP32 engine (int32 cell + FP32 offset positions, int64 fixed-point deposit, complex64
FFT) on the device (hgpm_copy_particles).  Both run the whole schedule and `compare`
returns (and `python -m hacc_gpm_amd.accuracy` prints as one JSON line)
  - pk_max_rel_err      : max relative P(k) error over bins with k < k_Nyq/2,
                          k_Nyq = pi NG / RL (fold 0), at the end of the run,
  - pk_counts_equal     : the two runs' mode counts agree in those bins,
  - rms_disp_cells      : RMS periodic position difference in grid units (FP64 sums on
                          the device, hgpm_compare_particles),
  - max_disp_cells      : the largest position difference of one component,
  - vel_rel_rms         : RMS velocity difference relative to the RMS FP64 velocity,
  - density_bit_exact   : two deposits of the final FP32 positions give the same int64 grid,
  - ms_per_step of both paths and their ratio.

    python -m hacc_gpm_amd.accuracy --ng 512 --rl 1000 --z-ini 50 --steps 100 --seed 1
    python -m hacc_gpm_amd.accuracy --ng 768 --np 384 --rl 1500 --z-ini 40 --steps 30 --seed 2000
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

import hacc_gpm_amd as hg


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run(m, sched, label, quiet=False):
    m.synchronize()
    t0 = time.perf_counter()
    for s, co in enumerate(sched):
        m.step(co)
        if (s + 1) % 10 == 0 or s + 1 == len(sched):
            m.synchronize()
            if not quiet:
                log(f"{label}: step {s + 1}/{len(sched)} {time.perf_counter() - t0:.1f} s")
    m.synchronize()
    return (time.perf_counter() - t0) * 1e3 / max(len(sched), 1)


def eh_table(p, k_max=100.0, k_bins=200001):
    """Eisenstein-Hu no-wiggle P(k) table of the run's cosmology at its z_ini, through the
    .ipk writer and reader (the reference's file format)."""
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "eh.ipk")
        hg.write_ipk_eh(p, fn, k_max, k_bins)
        return hg.read_ipk(fn)


def pk_compare(k, pk32, c32, pk64, c64, ng, rl):
    """Max relative P(k) error over the bins with k < k_Nyq/2 (k_Nyq = pi NG / RL, the
    line postprocessing/run.py:43-53 draws), and whether the mode counts agree there."""
    knyq = np.pi * ng / rl
    sel = (c64 > 0) & (k < knyq / 2) & (pk64 != 0)
    rel = np.abs(pk32[sel] - pk64[sel]) / np.abs(pk64[sel])
    return {"pk_bins_compared": int(sel.sum()), "pk_max_rel_err": float(rel.max()) if rel.size else None,
            "pk_counts_equal": bool(np.array_equal(c32[sel], c64[sel]))}


def compare(ng, np_, rl, z_ini, steps, seed, pk_bins=221, quiet=False, final=None):
    """The whole run z_ini -> 0 in `steps` steps on both paths from the FP64 path's initial
    conditions.  final(m32, m64), if given, is called with both contexts at the end of
    the run (before the two extra deposits) and its result is returned as "final"."""
    mk = lambda prec: hg.make_params(ng, np_, rl, z_ini, 0.0, steps, seed=seed, pk_bins=pk_bins, precision=prec)
    p64, p32 = mk(hg.FP64), mk(hg.FP32)
    ts = hg.Timestepper(p64)
    ts.setInitialZ(p64.z_ini)
    ts.reverseHalfStep()
    if not ts.aa > 0:
        raise ValueError("N_STEPS too small: a < 0 after the initial half step back")
    hdr, vals = eh_table(p64)
    out = {"workload": f"{np_}^3 particles, {ng}^3 grid, box {rl:g} Mpc/h, z={z_ini:g}->0, "
                       f"{steps} steps, seed {seed}"}
    with hg.MemoryManager(p64) as m64, hg.MemoryManager(p32) as m32:
        hg.GenerateDisplacementIC(m64, p64, ts, hdr, vals)
        m32.copy_particles_from(m64)
        ini = m32.compare_particles(m64)
        out["ini_max_disp_cells"] = ini["max_disp_cells"]
        if not quiet:
            log("initial conditions set in both contexts")
        out["ms_per_step_fp64"] = run(m64, hg.schedule(p64, steps, merge=False), "fp64", quiet)
        out["ms_per_step_fp32"] = run(m32, hg.schedule(p32, steps, merge=True), "fp32", quiet)
        out["speedup_fp32_vs_fp64"] = out["ms_per_step_fp64"] / out["ms_per_step_fp32"]
        k, pk64, c64 = m64.power_spectrum(0)
        _, pk32, c32 = m32.power_spectrum(0)
        out.update(pk_compare(k, pk32, c32, pk64, c64, ng, rl))
        d = m32.compare_particles(m64)
        out["rms_disp_cells"] = d["rms_disp_cells"]
        out["max_disp_cells"] = d["max_disp_cells"]
        out["vel_rel_rms"] = d["vel_rel_rms"]
        if final is not None:
            out["final"] = final(m32, m64)
        m32.cic()
        r1 = m32.get_density()
        m32.cic()
        out["density_bit_exact"] = bool(np.array_equal(r1, m32.get_density()))
        del r1
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ng", type=int, default=512)
    ap.add_argument("--np", type=int, default=0, help="particles per side (default: ng)")
    ap.add_argument("--rl", type=float, default=1000.0)
    ap.add_argument("--z-ini", type=float, default=50.0)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pk-bins", type=int, default=221)
    args = ap.parse_args()
    out = compare(args.ng, args.np or args.ng, args.rl, args.z_ini, args.steps, args.seed, args.pk_bins)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
