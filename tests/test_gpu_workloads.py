"""GPU tests at the benchmark workloads' full shapes.

  C2  512^3 particles on a 512^3 grid, box 1000 Mpc/h, z = 50 -> 0 in 100 steps, seed 1
  C3  1024^3 particles on a 1024^3 grid, box 2000 Mpc/h, z = 50 -> 0 in 50 steps, seed 2
  C4  uneven ensemble: 64^3, 128^3 and 192^3 boxes (grid = particles per side, 2 Mpc/h
      per cell), per-box seed and sigma_8 in [0.7, 0.9], z = 30 -> 0 in 40 steps
  C5  384^3 particles on 768^3 grids (box 1500 Mpc/h) mixed with 96^3 boxes (box
      192 Mpc/h), z = 40 -> 0 in 30 steps, the large boxes stored in waves

The FP32 engine runs each whole schedule beside the FP64 plain path from the same initial
conditions.  Bars: max relative P(k) error <= 1e-3 over the bins with k < k_Nyq/2 and
equal mode counts there (postprocessing/run.py:43-53 draws k_Nyq/2); RMS displacement
and velocity errors under the bounds stated in each test; the int64 density of the
final, clustered state bit-exact against the oracle's deposit (cic_pm.cu:46-87
restated) and across two deposits; ensembles bit-identical to single boxes and waves
bit-identical to resident batches.
"""
import os
import tempfile

import numpy as np
import pytest

from hacc_gpm_amd import accuracy

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def routes(m):
    return {k: m.get_tunable(k) for k in ("deposit_wide", "deposit_deferred", "deposit_deferred_full")}


@pytest.mark.timeout(600)
def test_c2_512_fp32_vs_fp64_whole_run(hg, orc):
    """C2 in full.  At z = 0 the final positions' deposit (64-bit windows and slab routes
    at full scale) is bit-exact against the oracle's int64 CIC."""
    ng = 512
    case(f"C2 {ng}^3 box 1000 z=50->0 100 steps")

    def final(m32, m64):
        cell, off, vel = m32.get_particles_split()
        del vel
        m32.cic()
        got = m32.get_density()
        r = routes(m32)
        ref = orc.cic_split(ng, cell, off)
        del cell, off
        return dict(r, bit_exact=bool(np.array_equal(got, ref)))

    out = accuracy.compare(ng, ng, 1000.0, 50.0, 100, 1, quiet=True, final=final)
    case(str(out))
    assert out["final"]["bit_exact"]
    assert out["density_bit_exact"]
    assert out["pk_counts_equal"] and out["pk_bins_compared"] > 50
    assert out["pk_max_rel_err"] <= 1e-3
    # measured 1.0e-4 cells and 7.8e-5: bounds with a factor 20 of headroom
    assert out["rms_disp_cells"] < 2e-3
    assert out["vel_rel_rms"] < 2e-3


@pytest.mark.timeout(900)
def test_c3_1024_fp32_vs_fp64_whole_run(hg, orc):
    """C3 in full: 1024^3 particles, 50 steps.  The final deposit is checked bit for bit
    against the oracle on a 96^3-cell block of the grid: the block's cells receive mass
    only from particles whose base cell lies in the block or one cell below it on each
    axis, so depositing exactly those particles (translated, on a grid that never wraps)
    gives the block's exact int64 values."""
    ng, S, x0 = 1024, 96, np.array([300, 517, 801])
    case(f"C3 {ng}^3 box 2000 z=50->0 50 steps")

    def final(m32, m64):
        cell, off, vel = m32.get_particles_split()
        del vel
        m32.cic()
        got = m32.get_density()[x0[0]:x0[0] + S, x0[1]:x0[1] + S, x0[2]:x0[2] + S].copy()
        r = routes(m32)
        sel = np.ones(cell.shape[0], bool)
        for d in range(3):
            sel &= (cell[:, d] >= x0[d] - 1) & (cell[:, d] <= x0[d] + S - 1)
        c = (cell[sel] - (x0 - 1)).astype(np.int32)
        o = off[sel]
        del cell, off
        ref = orc.cic_split(S + 2, c, o, 1.0)[1:S + 1, 1:S + 1, 1:S + 1]
        return dict(r, block_particles=int(sel.sum()), bit_exact=bool(np.array_equal(got, ref)),
                    block_mass=float(got.sum()) * 2.0 ** -24)

    out = accuracy.compare(ng, ng, 2000.0, 50.0, 50, 2, quiet=True, final=final)
    case(str(out))
    assert out["final"]["bit_exact"] and out["final"]["block_particles"] > S ** 3 // 2
    assert out["density_bit_exact"]
    assert out["pk_counts_equal"] and out["pk_bins_compared"] > 50
    assert out["pk_max_rel_err"] <= 1e-3
    # measured 1.3e-5 cells and 4.3e-6
    assert out["rms_disp_cells"] < 5e-4
    assert out["vel_rel_rms"] < 5e-4


def _ic(hg, p, k_max=100.0):
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "eh.ipk")
        hg.write_ipk_eh(p, fn, k_max, 20001)
        hdr, vals = hg.read_ipk(fn)
    a = 1.0 / (1.0 + p.z_ini)
    delta, _ = hg.growth(p, a)
    _, dot_delta = hg.growth(p, a - ts.deltaT / 2.0)
    return hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT


def _fp64_pk_of_box(hg, p, pos, vel, n_steps):
    """The FP64 plain path from a box's float3 initial state (grid units), whole schedule.
    A float3 position c + f rounds up to ng for a cell ng - 1 with f close to 1: it is
    wrapped to 0, the same point of the periodic box."""
    q = hg.make_params(p.ng, p.np, p.rl, p.z_ini, 0.0, p.nsteps, seed=p.seed, ss8=p.ss8,
                       pk_bins=p.pk_bins, precision=hg.FP64)
    x = pos.astype(np.float64)
    x[x >= p.ng] -= p.ng
    with hg.MemoryManager(q) as m:
        m.set_particles_f64(x, vel.astype(np.float64))
        for co in hg.schedule(q, n_steps, merge=False):
            m.step(co)
        return m.power_spectrum(0)


def _pk_check(hg, p, got, ref):
    k, pk32, c32 = got
    _, pk64, c64 = ref
    r = accuracy.pk_compare(k, pk32, c32, pk64, c64, p.ng, p.rl)
    assert r["pk_counts_equal"] and r["pk_bins_compared"] > 5, r
    assert r["pk_max_rel_err"] <= 1e-3, r
    return r


@pytest.mark.timeout(600)
def test_c4_uneven_ensemble_whole_run(hg):
    """C4's shapes: two boxes each at 64^3, 128^3 and 192^3 (2 Mpc/h per cell, seeds
    1000 + b, sigma_8 drawn uniformly in [0.7, 0.9] by a generator seeded 4242), the
    40 steps z = 30 -> 0 batched (the second half as one HIP graph run), against each
    box stepped alone: bit-identical particles and P(k).  One box of each class against
    the FP64 plain path from its initial state: P(k) within 1e-3."""
    rng = np.random.default_rng(4242)
    ps = []
    for n in (64, 128, 192):
        for _ in range(2):
            ps.append(hg.make_params(n, n, 2.0 * n, 30.0, 0.0, 40, seed=1000 + len(ps),
                                     ss8=float(rng.uniform(0.7, 0.9))))
    n_steps = 40
    case("C4 shapes: 2 x (64^3, 128^3, 192^3), 40 steps z=30->0")
    with hg.Ensemble(ps) as e:
        assert e.batches()[0] == 3
        ics = [_ic(hg, p) for p in ps]
        scheds = [hg.schedule(p, n_steps, merge=True) for p in ps]
        for b in range(len(ps)):
            e.generate_ic(b, *ics[b])
        ini = [e.get_particles(b) for b in range(len(ps))]
        for s in range(n_steps // 2):
            e.step([sc[s] for sc in scheds])
        e.run([[sc[s] for sc in scheds] for s in range(n_steps // 2, n_steps)])
        e.synchronize()
        for b, p in enumerate(ps):
            with hg.MemoryManager(p) as m:
                m.generate_ic(*ics[b])
                for s in range(n_steps):
                    m.step(scheds[b][s])
                pe, ve = e.get_particles(b)
                p1, v1 = m.get_particles()
                assert np.array_equal(pe, p1) and np.array_equal(ve, v1), b
                assert np.array_equal(e.power_spectrum(b)[1], m.power_spectrum(0)[1]), b
        for b in (0, 2, 4):
            r = _pk_check(hg, ps[b], e.power_spectrum(b), _fp64_pk_of_box(hg, ps[b], *ini[b], n_steps))
            case(f"box {b} ({ps[b].np}^3): {r}")


@pytest.mark.timeout(600)
def test_c5_pressure_waves_whole_run(hg):
    """C5's shapes: two boxes of 384^3 particles on 768^3 grids (box 1500 Mpc/h) and four
    of 96^3 (box 192 Mpc/h), z = 40 -> 0 in 30 steps.  Under a device budget that leaves
    grid buffers for one large box at a time the large batch steps in waves of one box;
    its particles and P(k) are bit-identical to the same ensemble fully resident.  One
    box of each class against the FP64 plain path from its initial state: P(k) within
    1e-3."""
    ps = [hg.make_params(768, 384, 1500.0, 40.0, 0.0, 30, seed=2000 + b) for b in range(2)]
    ps += [hg.make_params(96, 96, 192.0, 40.0, 0.0, 30, seed=2002 + b) for b in range(4)]
    n_steps = 30
    case("C5 shapes: 2 x 384^3 on 768^3 + 4 x 96^3, 30 steps z=40->0, waves vs resident")
    ics = {}
    for p in ps:
        if (p.ng, p.np) not in ics:
            ics[(p.ng, p.np)] = _ic(hg, p)
    scheds = [hg.schedule(p, n_steps, merge=True) for p in ps]
    with hg.Ensemble(ps) as res:
        nb, of = res.batches()
        w0 = res.waves()
        assert nb == 2 and int(w0[of[0]]) == 2 and int(w0[of[2]]) == 4
        full = res.device_bytes()
        for b, p in enumerate(ps):
            res.generate_ic(b, *ics[(p.ng, p.np)])
        ini = [res.get_particles(b) for b in (0, 2)]
        for s in range(n_steps):
            res.step([sc[s] for sc in scheds])
        res.synchronize()
        with hg.Ensemble(ps, device_budget=full - (4 << 30)) as wav:
            w = [int(x) for x in wav.waves()]
            big = int(of[0])
            case(f"resident {full / 1e9:.1f} GB; waves {w}, {wav.device_bytes() / 1e9:.1f} GB")
            assert w[big] == 1
            for b, p in enumerate(ps):
                wav.generate_ic(b, *ics[(p.ng, p.np)])
            for s in range(n_steps):
                wav.step([sc[s] for sc in scheds])
            for b in range(len(ps)):
                pr, vr = res.get_particles(b)
                pw, vw = wav.get_particles(b)
                assert np.array_equal(pr, pw) and np.array_equal(vr, vw), b
                assert np.array_equal(res.power_spectrum(b)[1], wav.power_spectrum(b)[1]), b
        for j, b in enumerate((0, 2)):
            r = _pk_check(hg, ps[b], res.power_spectrum(b), _fp64_pk_of_box(hg, ps[b], *ini[j], n_steps))
            case(f"box {b} ({ps[b].np}^3 on {ps[b].ng}^3): {r}")
