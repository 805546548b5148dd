"""Pins of the oracle functions that only the GPU comparison used to check: the fold of
P(k) (orc_fold_split / orc_fold_f64), the reference-rounding mode (orc_*_ref32 and the
REF32 branch of orc_solve_gradient), the FP64 kick and drift, and the composition of
the step loops (orc_run_*).  Each check is a closed form, an exact identity, brute
force or an independent library, chosen so that a plausible mistake (a wrong fold
direction, a dropped carry, a transposed operand, a wrong stage order) fails it."""
import itertools

import numpy as np
import pytest

from conftest import rand_split


# ------------------------------------------------------------------ fold ----
def test_fold_maps_x_to_twice_x_mod_half_box(orc):
    """foldParticles (fold_particles_power.cu:3-27): x/ng, minus 0.5 if >= 0.5, times
    2 ng, i.e. x -> 2 (x mod ng/2).  In the split layout it is exact: c + f is exact in
    FP64 for an FP32 offset, and so is twice it."""
    rng = np.random.default_rng(31)
    ng, n = 16, 20000
    cell, off = rand_split(rng, n, ng)
    c2, o2 = orc.fold_split(ng, cell, off)
    assert c2.min() >= 0 and c2.max() < ng and o2.min() >= 0 and o2.max() < 1
    x = cell.astype(np.float64) + off.astype(np.float64)
    want = 2.0 * np.mod(x, ng / 2)
    assert np.array_equal(c2.astype(np.float64) + o2.astype(np.float64), want)
    # FP64 form: the same map to rounding
    pos = rng.random((n, 3)) * ng
    p2 = orc.fold_f64(ng, pos)
    assert np.abs(p2 - 2.0 * np.mod(pos, ng / 2)).max() < 1e-12 * ng
    # hand cases: the half box maps to 0, just below it to just below ng, the carry of
    # an offset whose double reaches 1
    hand = [((0, 0.0), (0, 0.0)), ((8, 0.0), (0, 0.0)), ((7, 0.75), (15, 0.5)),
            ((12, 0.125), (8, 0.25)), ((15, 0.5), (15, 0.0)), ((3, 0.625), (7, 0.25))]
    for (c, f), (ce, fe) in hand:
        cc, oo = orc.fold_split(ng, np.array([[c, c, c]], np.int32), np.full((1, 3), f, np.float32))
        assert tuple(cc[0]) == (ce,) * 3 and tuple(oo[0]) == (np.float32(fe),) * 3, (c, f, cc, oo)
        pf = orc.fold_f64(ng, np.full((1, 3), c + f))
        assert np.allclose(pf, ce + fe, atol=1e-12)


def test_fold_of_a_half_periodic_density_doubles_its_modes(orc):
    """A density periodic in ng/2 (particles on lattice points, so CIC is exact) folds to
    8 copies on the even cells, and its Fourier modes become those of the unfolded
    density at twice the index: rho_f^(m) = rho^(2m mod ng).  Two folds of a density
    periodic in ng/4 give rho^(4m mod ng)."""
    rng = np.random.default_rng(32)
    ng = 16
    for folds, period in ((1, ng // 2), (2, ng // 4)):
        u = np.argwhere(rng.random((period,) * 3) < 0.3).astype(np.int32)
        shifts = np.array(list(itertools.product(range(0, ng, period), repeat=3)), np.int32)
        cell = (u[:, None, :] + shifts[None, :, :]).reshape(-1, 3)
        off = np.zeros(cell.shape, np.float32)
        rho0 = orc.cic_split(ng, cell, off)
        c, o = cell, off
        for _ in range(folds):
            c, o = orc.fold_split(ng, c, o)
        rhof = orc.cic_split(ng, c, o)
        s = 2 ** folds
        expect = np.zeros_like(rhof)
        expect[::s, ::s, ::s] = s ** 3 * rho0[:period, :period, :period]
        assert np.array_equal(rhof, expect)
        R0 = np.fft.fftn(rho0.astype(np.float64))
        Rf = np.fft.fftn(rhof.astype(np.float64))
        m = (np.arange(ng) * s) % ng
        np.testing.assert_allclose(Rf, R0[np.ix_(m, m, m)], rtol=0, atol=1e-9 * np.abs(R0).max())


# ---------------------------------------------------- reference FP32 mode ----
def test_cic_ref32_closed_forms_and_fp64_agreement(orc):
    ng = 8
    mass = 1.0
    pos = np.array([[0, 0, 0], [7, 7, 7], [3, 4, 5]], np.float32)
    rho = orc.cic_ref32(ng, pos, mass)
    assert rho[0, 0, 0] == 1 and rho[7, 7, 7] == 1 and rho[3, 4, 5] == 1 and rho.sum() == 3
    rho = orc.cic_ref32(ng, np.full((1, 3), 7.5, np.float32), 8.0)
    nz = np.argwhere(rho)
    assert set(map(tuple, nz)) == set(itertools.product((7, 0), repeat=3)) and np.all(rho[rho != 0] == 1.0)
    rng = np.random.default_rng(33)
    ng, n = 16, 20000
    pos = (rng.random((n, 3)) * ng).astype(np.float32)
    pos[pos >= ng] = 0
    r32 = orc.cic_ref32(ng, pos, mass)
    r64 = orc.cic_f64(ng, pos.astype(np.float64), mass)
    # FP32 weights and sums against FP64: a few FP32 ulps of each cell's total
    assert np.abs(r32 - r64).max() <= 64 * np.finfo(np.float32).eps * r64.max()
    assert abs(float(r32.astype(np.float64).sum()) - n) < 1e-4 * n
    # a fixed summation order: the same bits on every call
    assert np.array_equal(orc.cic_ref32(ng, pos, mass), r32)


def test_kick_ref32_and_f64_closed_forms(orc):
    rng = np.random.default_rng(34)
    ng, n = 8, 400
    pos = rng.random((n, 3)) * ng
    F = np.array([0.5, -0.25, 2.0])
    force = np.tile(F, (ng ** 3, 1))
    dt, fs = 0.01, 3.0
    v64 = orc.kick_f64(ng, pos, np.zeros((n, 3)), force, dt, fs)
    np.testing.assert_allclose(v64, np.tile(dt * fs * F, (n, 1)), rtol=1e-14)
    v32 = orc.kick_ref32(ng, pos.astype(np.float32), np.zeros((n, 3), np.float32), force.astype(np.float32), dt, fs)
    np.testing.assert_allclose(v32, np.tile(dt * fs * F, (n, 1)), rtol=4e-6)
    # a field linear in x is interpolated exactly (away from the periodic seam)
    x = np.arange(ng)
    Fx = np.zeros((ng, ng, ng, 3))
    Fx[..., 0] = x[:, None, None] * 0.125
    Fx[..., 2] = -x[:, None, None] * 0.5 + 1.0
    inside = pos[:, 0] < ng - 1
    v64 = orc.kick_f64(ng, pos, np.zeros((n, 3)), Fx.reshape(-1, 3), 1.0, 1.0)
    np.testing.assert_allclose(v64[inside, 0], 0.125 * pos[inside, 0], rtol=1e-13)
    np.testing.assert_allclose(v64[inside, 2], 1.0 - 0.5 * pos[inside, 0], rtol=1e-13, atol=1e-13)
    p32 = pos.astype(np.float32)
    v32 = orc.kick_ref32(ng, p32, np.zeros((n, 3), np.float32), Fx.reshape(-1, 3).astype(np.float32), 1.0, 1.0)
    np.testing.assert_allclose(v32[inside, 0], 0.125 * p32[inside, 0].astype(np.float64), rtol=4e-6, atol=4e-7)
    # random force grid: the two precisions agree to FP32 rounding
    force = rng.standard_normal((ng ** 3, 3))
    v64 = orc.kick_f64(ng, p32.astype(np.float64), np.zeros((n, 3)), force.astype(np.float32).astype(np.float64), dt, fs)
    v32 = orc.kick_ref32(ng, p32, np.zeros((n, 3), np.float32), force.astype(np.float32), dt, fs)
    assert np.abs(v32 - v64).max() < 2e-6 * np.abs(v64).max()


def test_drift_ref32_and_f64_closed_forms(orc):
    rng = np.random.default_rng(35)
    ng, n = 16, 5000
    pos = rng.random((n, 3)) * ng
    # moves under one box: UpdatePosKernel wraps with fmod(x + v p + ng, ng)
    # (update_positions_pm.cu:9-14), so a move of more than a box backwards stays
    # negative there, and the oracle follows it (pinned below)
    vel = np.clip(rng.normal(0, 15, (n, 3)), -40, 40)
    pf = 0.37
    p64 = orc.drift_f64(ng, pos, vel, pf)
    assert p64.min() >= 0 and p64.max() < ng
    far = orc.drift_f64(ng, np.array([[1.0, 1.0, 1.0]]), np.array([[-20.0, 0.0, 0.0]]), 1.0)
    assert far[0, 0] == np.fmod(1.0 - 20.0 + ng, ng) and far[0, 0] < 0
    np.testing.assert_allclose(p64, np.mod(pos + vel * pf, ng), rtol=0, atol=1e-12)
    p32 = pos.astype(np.float32)
    v32 = vel.astype(np.float32)
    q32 = orc.drift_ref32(ng, p32, v32, np.float32(pf))
    assert q32.min() >= 0 and q32.max() <= ng
    want = np.mod(p32.astype(np.float64) + v32.astype(np.float64) * np.float64(np.float32(pf)), ng)
    d = (q32.astype(np.float64) - want + ng / 2) % ng - ng / 2
    assert np.abs(d).max() < 8 * np.finfo(np.float32).eps * ng
    # a move that crosses the box face wraps
    q = orc.drift_ref32(ng, np.array([[15.5, 0.25, 8.0]], np.float32), np.array([[1.0, -1.0, 0.0]], np.float32),
                        np.float32(1.0))
    np.testing.assert_array_equal(q, [[0.5, 15.25, 8.0]])


def test_solve_ref32_single_mode_closed_form(orc):
    """The REF32 branch (FP32 grid data, float cos/sin of float k, FP64 sum and divide)
    gives the single-mode closed form to FP32 accuracy, and the F64 mode's forces to
    FP32 rounding on a random density."""
    ng = 16
    m = np.array([1, 2, 3])
    I = np.indices((ng, ng, ng))
    phase = 2 * np.pi * (m[0] * I[0] + m[1] * I[1] + m[2] * I[2]) / ng
    rho = 1.0 + 0.3 * np.cos(phase)
    F = orc.solve_gradient(ng, rho, orc.REF32).reshape(ng, ng, ng, 3)
    k = 2 * np.pi * m / ng
    lam = 2 * np.sum(np.cos(k) - 1)
    for d in range(3):
        expect = 0.3 * np.sin(k[d]) * np.sin(phase) / lam
        np.testing.assert_allclose(F[..., d], expect, atol=2e-6 * np.abs(expect).max())
    rng = np.random.default_rng(36)
    rho = 1 + 0.3 * rng.standard_normal((ng, ng, ng))
    F32 = orc.solve_gradient(ng, rho, orc.REF32)
    F64 = orc.solve_gradient(ng, rho, orc.F64)
    assert np.sqrt(((F32 - F64) ** 2).mean() / (F64 ** 2).mean()) < 2e-6
    assert np.array_equal(F32, F32.astype(np.float32).astype(np.float64))  # stored as FP32


# ------------------------------------------------------ step loop order ----
def _ic(orc, hg, ng, seed):
    import os
    import tempfile
    p = hg.make_params(ng, ng, 2.0 * ng, 30.0, 0.0, 20, seed=seed)
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "eh.ipk")
        hg.write_ipk_eh(p, fn, 20.0, 20001)
        hdr, vals = hg.read_ipk(fn)
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    a = 1 / (1 + p.z_ini)
    D, _ = orc.growth_f64(a, p.omega_matter, 1 - p.omega_matter)
    _, dD = orc.growth_f64(a - ts.deltaT / 2, p.omega_matter, 1 - p.omega_matter)
    pos, vel = orc.ic_f64(ng, ng, p.rl, seed, vals, hdr[3], D, dD, p.z_ini, ts.fscal, ts.deltaT)
    return p, pos, vel, hg.schedule_array(p, 3, merge=False)


def test_run_loops_are_the_serial_step_order(orc, hg):
    """orc_run_* step the serial() loop body (driver_pm.cpp:110-125): drift, deposit,
    solve, kick, drift.  Composing the (separately pinned) primitives in that order
    gives the same bits for each mode; a swapped order (kick before the deposit, the
    drifts' coefficients exchanged) does not."""
    ng = 16
    p, pos, vel, co = _ic(orc, hg, ng, 41)
    st = {}
    p64, v64, _ = orc.run_f64(ng, ng, co, pos, vel, stages=st)
    assert set(st) == set(orc.STAGES) and all(v >= 0 for v in st.values())
    x, v = pos.copy(), vel.copy()
    for c in co:
        x = orc.drift_f64(ng, x, v, c[0])
        F = orc.solve_gradient(ng, orc.cic_f64(ng, x))
        v = orc.kick_f64(ng, x, v, F, c[1], c[2])
        x = orc.drift_f64(ng, x, v, c[3])
    assert np.array_equal(x, p64) and np.array_equal(v, v64)
    # the drifts' coefficients exchanged: a different state
    xs, vs = pos.copy(), vel.copy()
    for c in co:
        xs = orc.drift_f64(ng, xs, vs, c[3])
        F = orc.solve_gradient(ng, orc.cic_f64(ng, xs))
        vs = orc.kick_f64(ng, xs, vs, F, c[1], c[2])
        xs = orc.drift_f64(ng, xs, vs, c[0])
    assert not np.array_equal(vs, v64)
    # split mode: int64 density * 2^-24, FP64 solve rounded to FP32 forces
    cell, off = orc.split_positions(ng, pos)
    c2, o2, v2, _ = orc.run_split(ng, ng, co, cell, off, vel.astype(np.float32))
    c, o, v = cell, off, vel.astype(np.float32)
    for k in co:
        c, o = orc.drift_split(ng, c, o, v, np.float32(k[0]))
        rho = orc.cic_split(ng, c, o) * 2.0 ** -24
        F = orc.solve_gradient(ng, rho).astype(np.float32)
        v = orc.kick_split(ng, c, o, v, F, k[1], k[2])
        c, o = orc.drift_split(ng, c, o, v, np.float32(k[3]))
    assert np.array_equal(c, c2) and np.array_equal(o, o2) and np.array_equal(v, v2)
    # reference FP32 mode
    p3, v3, _ = orc.run_ref32(ng, ng, co, pos.astype(np.float32), vel.astype(np.float32))
    x, v = pos.astype(np.float32), vel.astype(np.float32)
    for k in co:
        x = orc.drift_ref32(ng, x, v, np.float32(k[0]))
        F = orc.solve_gradient(ng, orc.cic_ref32(ng, x, 1.0).astype(np.float64), orc.REF32).astype(np.float32)
        v = orc.kick_ref32(ng, x, v, F, k[1], k[2])
        x = orc.drift_ref32(ng, x, v, np.float32(k[3]))
    assert np.array_equal(x, p3) and np.array_equal(v, v3)
    # and the three modes agree to FP32 accuracy
    d = (c2 + o2.astype(np.float64) - p64 + ng / 2) % ng - ng / 2
    assert np.sqrt((d ** 2).mean()) < 1e-4
    d = (p3.astype(np.float64) - p64 + ng / 2) % ng - ng / 2
    assert np.sqrt((d ** 2).mean()) < 1e-4


def test_kick_split_against_f64(orc):
    """The build's FP32 kick (S = sum w F by fmaf, then v = fmaf(float(dt fscal), S, v))
    against the FP64 kick of the same positions and forces: within FP32 rounding of the
    8-term sum, for random fields and velocities."""
    rng = np.random.default_rng(31)
    ng, n = 12, 4000
    cell, off = rand_split(rng, n, ng)
    pos = cell.astype(np.float64) + off.astype(np.float64)
    F = rng.normal(0, 1, (ng ** 3, 3)).astype(np.float32)
    v0 = rng.normal(0, 1, (n, 3)).astype(np.float32)
    dt, fs = 0.0137, 2.71
    v32 = orc.kick_split(ng, cell, off, v0, F, dt, fs).astype(np.float64)
    v64 = orc.kick_f64(ng, pos, v0.astype(np.float64), F.astype(np.float64), dt, fs)
    dv = v64 - v0
    err = np.abs(v32 - v64)
    assert err.max() <= 4e-7 * (np.abs(v0).max() + np.abs(dv).max())
    # the kick is not the identity: a dropped factor or a lost corner fails here
    assert np.abs(dv).max() > 1e-3


def test_drift_split_fixed_point(orc):
    """The split drift moves a fixed-point coordinate (2^-22 cells) by the FP32
    displacement v * p rounded to the nearest 2^-22 cell.  Pins: a displacement that is
    a multiple of 2^-22 moves exactly by it; ties go to even (a half step does not move,
    one and a half steps move two); a drift by -t undoes a drift by +t exactly; moves
    wrap modulo ng; and any drift agrees with the FP64 drift of the same displacement
    to within half a step."""
    ng = 16
    u = 2.0 ** -22
    cell = np.array([[3, 15, 0]], np.int32)
    off = np.array([[0.25, 1.0 - u, 0.0]], np.float32)
    one = np.ones((1, 3), np.float32)

    def drift(c, o, t):
        return orc.drift_split(ng, c, o, one * np.float32(t), np.float32(1.0))

    # exact multiples of the step, across the periodic edge in y and below 0 in z
    c, o = drift(cell, off, 5 * u)
    x = c + o.astype(np.float64)
    assert np.array_equal(x, np.array([[3.25 + 5 * u, (5 - 1) * u, 5 * u]]))
    c, o = drift(cell, off, -3 * u)
    x = c + o.astype(np.float64)
    assert np.array_equal(x, np.array([[3.25 - 3 * u, 16 - 4 * u, 16 - 3 * u]]))
    # ties to even: 0.5 step -> 0, 1.5 steps -> 2, 2.5 steps -> 2, -0.5 step -> 0
    c3, o3 = np.full((1, 3), 3, np.int32), np.full((1, 3), 0.25, np.float32)
    for t, k in ((0.5, 0), (1.5, 2), (2.5, 2), (-0.5, 0), (-1.5, -2)):
        c, o = drift(c3, o3, t * u)
        assert np.all(c + o.astype(np.float64) == 3.25 + k * u), (t, c, o)
    # whole cells plus a fraction, and more than a box (reduced modulo ng)
    c, o = drift(cell, off, 7.75)
    assert np.array_equal(c + o.astype(np.float64), np.array([[11.0, 7.75 - u, 7.75]]))
    c, o = drift(cell, off, 16.0 + 2.0)
    assert np.array_equal(c, np.array([[5, 1, 2]])) and np.array_equal(o, np.array([[0.25, 1 - u, 0.0]], np.float32))
    # reversible, and within half a step of the FP64 drift
    rng = np.random.default_rng(41)
    cs, os_ = rand_split(rng, 5000, ng)
    vel = rng.normal(0, 3, (5000, 3)).astype(np.float32)
    for pf in (0.37, 1.9, 0.013):
        c1, o1 = orc.drift_split(ng, cs, os_, vel, np.float32(pf))
        c2, o2 = orc.drift_split(ng, c1, o1, -vel, np.float32(pf))
        assert np.array_equal(c2, cs) and np.array_equal(o2, os_)
        t = (vel * np.float32(pf)).astype(np.float64)
        want = (cs + os_.astype(np.float64) + t) % ng
        d = (c1 + o1.astype(np.float64) - want + ng / 2) % ng - ng / 2
        assert np.abs(d).max() <= u / 2 * (1 + 1e-9)
