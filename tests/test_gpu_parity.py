"""GPU parity tests: the HIP engine (through the C ABI of libhgpm.so) against the CPU
oracle on the same inputs.

Bars:
  - int64 density, cell indices, offsets and velocities given the same force grid:
    bit-exact (integer / FP32 recipes without FMA contraction on both sides);
  - FFT: relative L2 error vs FP64 <= 2e-6 (FP32) / 1e-13 (FP64);
  - forces (FP32 solve) vs FP64 oracle: relative RMS <= 2e-5;
  - P(k): counts exact, values relative <= 1e-4 (FP32 spectrum);
  - full runs: stated in each test.
"""
import os
import sys

import numpy as np
import pytest

from conftest import lattice_split, rand_split, split_of

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


def mk(hg, ng, np_=None, rl=None, prec=0, **kw):
    np_ = np_ or ng
    rl = rl or 2.0 * ng
    return hg.MemoryManager(hg.make_params(ng, np_, rl, 50.0, 0.0, 100, precision=prec, **kw))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


# --------------------------------------------------------------- deposit ----
@pytest.mark.parametrize("ng,clustered", [(32, False), (64, False), (64, True), (96, True)])
def test_deposit_bit_exact(hg, orc, ng, clustered):
    case(f"deposit ng={ng} clustered={clustered}")
    rng = np.random.default_rng(ng + clustered)
    with mk(hg, ng) as m:
        cell, off = rand_split(rng, m.N, ng, clustered)
        # particles on the periodic seam (one position step below the box edge) and
        # exactly on lattice points
        cell[:5] = ng - 1
        off[:5] = np.float32(1.0) - np.float32(2.0 ** -22)
        off[5:9] = 0.0
        vel = np.zeros((m.N, 3), np.float32)
        m.set_particles_split(cell, off, vel)
        m.cic()
        got = m.get_density()
        ref = orc.cic_split(ng, cell, off)
        assert np.array_equal(got, ref)
        m.cic()  # deterministic across runs
        assert np.array_equal(m.get_density(), got)


@pytest.mark.parametrize("ng,np_,clustered", [(96, 64, False), (96, 64, True), (64, 32, True)])
def test_deposit_bit_exact_particle_mass(hg, orc, ng, np_, clustered):
    """Particle mass (ng/np)^3 other than 1: 3.375 (not a power of two: the weights'
    product with the mass rounds) and 8 (a power of two: the windows pass folds the
    mass into the z factors, an exact scaling).  Bit-exact against the oracle."""
    case(f"deposit mass ng={ng} np={np_} clustered={clustered}")
    rng = np.random.default_rng(7 * ng + np_ + clustered)
    gp = np.float32(ng) / np.float32(np_)
    mass = float(gp * gp * gp)
    with mk(hg, ng, np_) as m:
        cell, off = rand_split(rng, m.N, ng, clustered)
        vel = np.zeros((m.N, 3), np.float32)
        m.set_particles_split(cell, off, vel)
        m.cic()
        got = m.get_density()
        ref = orc.cic_split(ng, cell, off, mass)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("ng,np_", [(64, 64), (128, 64), (256, 256)])
def test_deposit_routes(hg, orc, ng, np_):
    """Arena route (windows -> owner blocks), atomic route for deferred tiles, and the
    mixes forced by a short arena or a short slot list: all bit-identical.  256^3: 4096
    tiles over 16 sub-arenas."""
    case(f"deposit routes ng={ng} np={np_}")
    rng = np.random.default_rng(ng * 7 + np_)
    mass = (ng / np_) ** 3
    with mk(hg, ng, np_) as m:
        cell, off = lattice_split(rng, np_, ng, 1.5, outliers=3)
        m.set_particles_split(cell, off, np.zeros((m.N, 3), np.float32))
        ref = orc.cic_split(ng, cell, off, mass)
        ntiles = m.N // 4096  # deposit tiles: one 16^3 Lagrangian tile
        m.cic()
        assert np.array_equal(m.get_density(), ref)
        d0 = m.get_tunable("deposit_deferred")
        assert d0 < ntiles  # most tiles took the arena route
        slots0 = m.get_tunable("deposit_slots")
        for key, val in [("deposit_arena", 0), ("deposit_arena", 3 * 9000), ("deposit_slots", 1),
                         ("deposit_slots", 4)]:
            m.set_tunable(key, val)
            m.cic()
            assert np.array_equal(m.get_density(), ref), (key, val)
            d = m.get_tunable("deposit_deferred")
            if key == "deposit_arena" and val == 0:
                assert d == ntiles
            elif val != 4:  # a 4-slot list may or may not overflow
                assert 0 < d <= ntiles, (key, val, d)
            m.set_tunable("deposit_arena", -1)
            m.set_tunable("deposit_slots", slots0)
        m.cic()
        assert np.array_equal(m.get_density(), ref)
        # which tiles find arena room first depends on scheduling: only the bits are fixed
        assert m.get_tunable("deposit_deferred") < ntiles


@pytest.mark.parametrize("np_", [64, 32])
def test_deposit_dense_cells_take_the_wide_route(hg, orc, np_):
    """Tiles whose window holds a cell above 2^32 - 1 mass units (hundreds of particles
    of one tile in one cell) store their window again with 64-bit cells: nothing is
    deferred and the density stays bit-exact.  With no arena room left they take the
    atomic route instead.  A step over such tiles keeps the FP32 density handoff and
    matches the step made of separate calls bit for bit."""
    ng = 64
    case(f"deposit dense cells ng={ng} np={np_}")
    rng = np.random.default_rng(77 + np_)
    p = hg.make_params(ng, np_, 2.0 * ng, 50.0, 0.0, 100)
    with hg.MemoryManager(p) as m, hg.MemoryManager(p) as m2:
        x = (np.array([20.0, 33.0, 41.0]) + rng.normal(0, 0.3, (m.N, 3))) % ng
        cell, off = split_of(x, ng)
        vel = np.zeros((m.N, 3), np.float32)  # at rest: the first step deposits the clump
        ref = orc.cic_split(ng, cell, off, (ng / np_) ** 3)
        m.set_particles_split(cell, off, vel)
        m.cic()
        assert np.array_equal(m.get_density(), ref)
        assert m.get_tunable("deposit_wide") > 0
        assert m.get_tunable("deposit_deferred") == 0
        assert m.get_tunable("deposit_owner_redo") > 0  # 64-bit chunks: the 64-bit owner pass
        m.set_tunable("deposit_arena", 100 * (m.N // 4096))  # room for part of the windows only
        m.cic()
        assert np.array_equal(m.get_density(), ref)
        assert m.get_tunable("deposit_deferred_full") > 0
        m.set_tunable("deposit_arena", -1)
        m2.set_particles_split(cell, off, vel)
        for s, co in enumerate(hg.schedule(p, 3, merge=True)):
            m.step(co)
            if s == 0:
                assert m.get_tunable("deposit_wide") > 0 and m.get_tunable("deposit_deferred") == 0
            if not co.skip_drift0:
                m2.update_positions(co.drift0)
            m2.cic()
            m2.solve_gradient()
            m2.kick_drift(co.kick_dt, co.kick_fscal, co.drift1)
            for a, b in zip(m.get_particles_split(), m2.get_particles_split()):
                assert np.array_equal(a, b)


def test_deposit_owner_sum_wraps_32_bits(hg, orc):
    """The owner pass sums chunks in 32-bit LDS cells and hands a block to the 64-bit
    owner pass when a cell's sum reaches 2^32 units.  Eight Lagrangian tiles each pile
    208 particles onto one lattice point (3.5e9 units per tile: no tile window holds a
    cell over 32 bits, so no 64-bit window), 2.8e10 units together: the block is
    redone in 64 bits and the density is bit-exact against the oracle; the lattice
    alone needs no redo."""
    ng = 64
    case(f"deposit owner 32-bit wrap ng={ng}")
    rng = np.random.default_rng(5)
    p = hg.make_params(ng, ng, 2.0 * ng, 50.0, 0.0, 100)
    with hg.MemoryManager(p) as m:
        cell, off = lattice_split(rng, ng, ng, 0.3)
        m.set_particles_split(cell, off, np.zeros((m.N, 3), np.float32))
        m.cic()
        assert np.array_equal(m.get_density(), orc.cic_split(ng, cell, off))
        assert m.get_tunable("deposit_owner_redo") == 0
        q = np.indices((ng,) * 3).reshape(3, -1).T
        pile = (q[:, 0] % 16 < 13) & (q[:, 1] % 16 == 0) & (q[:, 0] < 32) & (q[:, 1] < 32) & (q[:, 2] < 32)
        assert pile.sum() == 8 * 13 * 16
        cell[pile] = 30
        off[pile] = 0.0
        m.set_particles_split(cell, off, np.zeros((m.N, 3), np.float32))
        m.cic()
        got = m.get_density()
        assert np.array_equal(got, orc.cic_split(ng, cell, off))
        assert got.max() >= 2 ** 32
        assert m.get_tunable("deposit_owner_redo") > 0 and m.get_tunable("deposit_wide") == 0
        # the same through the step's FP32 density handoff: one step matches the
        # separate calls bit for bit
        with hg.MemoryManager(p) as m2:
            vel = rng.normal(0, 0.3, (m.N, 3)).astype(np.float32)
            m.set_particles_split(cell, off, vel)
            m2.set_particles_split(cell, off, vel)
            co = hg.schedule(p, 2, merge=True)[0]
            m.step(co)
            if not co.skip_drift0:
                m2.update_positions(co.drift0)
            m2.cic()
            m2.solve_gradient()
            m2.kick_drift(co.kick_dt, co.kick_fscal, co.drift1)
            for a, b in zip(m.get_particles_split(), m2.get_particles_split()):
                assert np.array_equal(a, b)

def test_deposit_mass_ratio(hg, orc):
    case("deposit np=32 ng=64 mass 8")
    rng = np.random.default_rng(1)
    with mk(hg, 64, 32) as m:
        cell, off = rand_split(rng, m.N, 64)
        m.set_particles_split(cell, off, np.zeros((m.N, 3), np.float32))
        m.cic()
        assert np.array_equal(m.get_density(), orc.cic_split(64, cell, off, 8.0))


@pytest.mark.parametrize("ng,np_", [(64, 64), (128, 64), (256, 256)])
def test_deposit_half_tiles(hg, orc, ng, np_):
    """The deposit with half Lagrangian tiles per workgroup (tunable deposit_tile 1: 8 x
    16 x 16 lattice blocks, twice the tiles): bit-exact against the oracle on the arena
    and on the atomic route, and a step with it (the gather staging from the half tiles'
    boxes) bit-identical to a step with whole tiles and to the separate calls."""
    case(f"deposit half tiles ng={ng} np={np_}")
    rng = np.random.default_rng(ng * 3 + np_)
    mass = (ng / np_) ** 3
    p = hg.make_params(ng, np_, 2.0 * ng, 50.0, 0.0, 100)
    with hg.MemoryManager(p) as m, hg.MemoryManager(p) as m2, hg.MemoryManager(p) as m3:
        cell, off = lattice_split(rng, np_, ng, 1.5, outliers=3)
        vel = rng.normal(0, 0.7, (m.N, 3)).astype(np.float32)
        ref = orc.cic_split(ng, cell, off, mass)
        m.set_tunable("deposit_tile", 1)
        assert m.get_tunable("deposit_tile") == 1
        m.set_particles_split(cell, off, vel)
        m.cic()
        assert np.array_equal(m.get_density(), ref)
        assert m.get_tunable("deposit_deferred") < m.N // 2048
        m.set_tunable("deposit_arena", 0)
        m.cic()
        assert np.array_equal(m.get_density(), ref)
        assert m.get_tunable("deposit_deferred") == m.N // 2048
        m.set_tunable("deposit_arena", -1)
        m2.set_particles_split(cell, off, vel)
        m3.set_tunable("deposit_tile", 1)
        m3.set_particles_split(cell, off, vel)
        for co in hg.schedule(p, 3, merge=True):
            m.step(co)
            m2.step(co)
            if not co.skip_drift0:
                m3.update_positions(co.drift0)
            m3.cic()
            m3.solve_gradient()
            m3.kick_drift(co.kick_dt, co.kick_fscal, co.drift1)
            a, b, c = m.get_particles_split(), m2.get_particles_split(), m3.get_particles_split()
            for x, y, z in zip(a, b, c):
                assert np.array_equal(x, y) and np.array_equal(x, z)


def test_autotune_particles_leaves_particles(hg):
    """hgpm_autotune_particles times the deposit tile and gather shapes on the context's
    particles and keeps valid choices; positions and velocities are unchanged."""
    case("autotune particles 64^3")
    p = hg.make_params(64, 64, 128.0, 50.0, 0.0, 100, seed=9)
    with hg.MemoryManager(p) as m:
        with pytest.raises(hg.HgpmError) as e:  # no particles yet: refused
            m._c(hg.lib().hgpm_autotune_particles(m.h, 1, None, None))
        assert e.value.code == hg.ERR_STATE
        rng = np.random.default_rng(9)
        cell, off = lattice_split(rng, 64, 64, 2.0)
        vel = rng.normal(0, 1, (m.N, 3)).astype(np.float32)
        vel[:10] = -0.0
        m.set_particles_split(cell, off, vel)
        t = m.autotune(2)
        assert t["deposit_tile"] in (0, 1) and t["gather_tile"] in (0, 1)
        assert all(x > 0 for x in t["deposit_ms"] + t["gather_ms"])
        c2, o2, v2 = m.get_particles_split()
        assert np.array_equal(c2, cell) and np.array_equal(o2, off)
        assert np.array_equal(v2.view(np.uint32), vel.view(np.uint32))


@pytest.mark.parametrize("ng,np_,sigma", [(192, 32, 1.5), (192, 32, 0.2), (256, 32, 1.5)])
def test_deposit_heavy_particles(hg, orc, ng, np_, sigma):
    """Particles of mass (ng/np)^3 >= 128 (216 at 192/32, 512 at 256/32) take the
    deposit kernels built for heavy particles (64-bit weights per particle).  A 16^3
    Lagrangian tile spans 96-128 cells per side; at mass 512 one particle's weight can
    exceed 2^32 - 1 units, and with sigma 0.2 a clump piles thousands of particles into
    a few cells (the 64-bit pass), then every tile on the atomic route under an empty
    arena.  All bit-exact against the oracle, and a step over these particles matches
    the separate calls bit for bit."""
    case(f"deposit heavy ng={ng} np={np_} sigma={sigma}")
    rng = np.random.default_rng(ng + np_ + int(10 * sigma))
    mass = (ng / np_) ** 3
    p = hg.make_params(ng, np_, 2.0 * ng, 50.0, 0.0, 100)
    with hg.MemoryManager(p) as m, hg.MemoryManager(p) as m2:
        cell, off = lattice_split(rng, np_, ng, sigma, outliers=2)
        if sigma < 1:  # a clump: 4096 particles piled into a few cells
            x = (np.array([50.0, 70.0, 90.0]) + rng.normal(0, 0.4, (4096, 3))) % ng
            cell[:4096], off[:4096] = split_of(x, ng)
        vel = rng.normal(0, 0.5, (m.N, 3)).astype(np.float32)
        ref = orc.cic_split(ng, cell, off, mass)
        m.set_particles_split(cell, off, vel)
        m.cic()
        assert np.array_equal(m.get_density(), ref)
        wide, deferred = m.get_tunable("deposit_wide"), m.get_tunable("deposit_deferred")
        case(f"  routes: wide {wide} deferred {deferred}")
        # a tile's window spans 96-128 cells per side: those over 512 owner blocks take
        # the atomic route (every tile at 256/32), heavy cells the 64-bit pass
        assert wide + deferred > 0
        m.set_tunable("deposit_arena", 0)  # every tile on the atomic route
        m.cic()
        assert np.array_equal(m.get_density(), ref)
        assert m.get_tunable("deposit_deferred") == m.N // 4096
        m.set_tunable("deposit_arena", -1)
        m2.set_particles_split(cell, off, vel)
        for co in hg.schedule(p, 3, merge=True):
            m.step(co)
            if not co.skip_drift0:
                m2.update_positions(co.drift0)
            m2.cic()
            m2.solve_gradient()
            m2.kick_drift(co.kick_dt, co.kick_fscal, co.drift1)
            for a, b in zip(m.get_particles_split(), m2.get_particles_split()):
                assert np.array_equal(a, b)


# ------------------------------------------------------------------- FFT ----
@pytest.mark.parametrize("ng", [32, 64, 96, 128, 192])
def test_fft_fp32(hg, ng):
    case(f"fft fp32 ng={ng}")
    rng = np.random.default_rng(ng)
    x = (rng.standard_normal((ng,) * 3) + 1j * rng.standard_normal((ng,) * 3)).astype(np.complex64)
    with mk(hg, ng) as m:
        X = m.fft(x, -1)
        ref = np.fft.fftn(x.astype(np.complex128))
        assert rel_l2(X, ref) < 2e-6
        back = m.fft(X, +1)
        assert rel_l2(back, ng ** 3 * x.astype(np.complex128)) < 4e-6


@pytest.mark.parametrize("ng,plan", [(256, 0), (384, 0), (512, 0), (768, 0), (1024, 0), (1024, 1)])
def test_fft_fp32_separable(hg, ng, plan):
    """The larger plans (256 = 16*16, 384 = 3*8*16, 512 = 8*8*8, 768 = 3*16*16,
    1024 = 16*8*8 and, "fft_plan" 1, 8*8*16): a separable grid a(x) b(y) c(z)
    transforms to A(kx) B(ky) C(kz), with A, B, C the 1-D transforms (numpy, FP64),
    checked at sampled modes; the inverse returns ng^3 times the grid."""
    case(f"fft fp32 separable ng={ng} plan={plan}")
    rng = np.random.default_rng(ng + 5)
    v = [(rng.standard_normal(ng) + 1j * rng.standard_normal(ng)).astype(np.complex64) for _ in range(3)]
    x = (v[0][:, None] * v[1][None, :])[:, :, None] * v[2][None, None, :]
    V = [np.fft.fft(u.astype(np.complex128)) for u in v]
    idx = rng.integers(0, ng, (200000, 3))
    with mk(hg, ng, 32) as m:
        if plan:
            m.set_tunable("fft_plan", plan)
        X = m.fft(x, -1)
        got = X[idx[:, 0], idx[:, 1], idx[:, 2]].astype(np.complex128)
        want = V[0][idx[:, 0]] * V[1][idx[:, 1]] * V[2][idx[:, 2]]
        assert rel_l2(got, want) < 5e-6
        back = m.fft(X, +1)
        del X
        got = back[idx[:, 0], idx[:, 1], idx[:, 2]].astype(np.complex128)
        want = float(ng) ** 3 * x[idx[:, 0], idx[:, 1], idx[:, 2]].astype(np.complex128)
        assert rel_l2(got, want) < 1e-5


@pytest.mark.parametrize("ng", [32, 64, 96])
def test_fft_fp64(hg, ng):
    case(f"fft fp64 ng={ng}")
    rng = np.random.default_rng(ng + 7)
    x = rng.standard_normal((ng,) * 3) + 1j * rng.standard_normal((ng,) * 3)
    with mk(hg, ng, prec=1) as m:
        X = m.fft(x, -1)
        assert rel_l2(X, np.fft.fftn(x)) < 1e-13
        assert rel_l2(m.fft(X, +1), ng ** 3 * x) < 1e-13


# ----------------------------------------------------------------- solve ----
@pytest.mark.parametrize("ng", [32, 64, 96])
def test_solve_fp32_vs_fp64_oracle(hg, orc, ng):
    case(f"solve fp32 ng={ng}")
    rng = np.random.default_rng(ng + 3)
    with mk(hg, ng) as m:
        cell, off = rand_split(rng, m.N, ng, clustered=False)
        # add structure: pull particles toward a few centres
        m.set_particles_split(cell, off, np.zeros((m.N, 3), np.float32))
        m.cic()
        q = m.get_density()
        m.solve_gradient()
        F = m.get_force().astype(np.float64)
    Fo = orc.solve_gradient(ng, q * 2.0 ** -24)
    err = np.sqrt(((F - Fo) ** 2).mean()) / np.sqrt((Fo ** 2).mean())
    assert err < 2e-5, err


@pytest.mark.parametrize("ng", [256, 512, 768, 1024])
def test_solve_fp32_vs_fp64_path_large(hg, ng):
    """The FP32 potential solve at the larger plans (half spectra: forward z, y, fused
    x pass, inverse y, inverse z; forces as central differences of the potential)
    against the FP64 plain path's k-space gradient of the same deposit: rel L2 < 2e-5,
    the bound the FP64 oracle check uses at small sizes."""
    case(f"solve fp32 vs fp64 path ng={ng}")
    rng = np.random.default_rng(ng + 9)
    np_ = ng // 2
    with mk(hg, ng, np_) as m:
        cell, off = lattice_split(rng, np_, ng, 2.0)
        m.set_particles_split(cell, off, np.zeros((np_ ** 3, 3), np.float32))
        del cell, off
        m.cic()
        q = m.get_density()
        m.solve_gradient()
        F = m.get_force()
    with mk(hg, ng, 32, prec=1) as m64:
        m64.set_density_f64(q.astype(np.float64) * 2.0 ** -24)
        del q
        m64.solve_gradient()
        F64 = m64.get_force()
    assert rel_l2(F.astype(np.float64), F64) < 2e-5


@pytest.mark.parametrize("ng,plan,xs", [(512, 0, -1), (768, 0, -1), (1024, 0, -1), (1024, 1, -1), (1024, 1, 5)])
def test_solve_single_mode_closed_form_large(hg, ng, plan, xs):
    """The FP32 half-spectrum solve at the large plans against the closed form, with no
    other code path in between: rho = 1 + A cos(k.x) gives, under the reference's
    Green's function (calcGreens.hpp:11-26) and central-difference gradient
    (kspace_solver.cu:55-90), F_d = A sin(k_d) sin(k.x) / (2 sum_d (cos k_d - 1)).
    The density goes in as int64 2^-24 units (rounding error 2^-25, relative 1e-7 of A).
    Relative RMS error over 2^21 sampled cells < 2e-5.  At 1024 also with the second
    radix order ("fft_plan" 1), with the default x pass and with 16 x 1024."""
    case(f"solve single mode closed form ng={ng} plan={plan} xs={xs}")
    A = 0.3
    m_ = np.array([ng // 8 + 1, 5, ng // 3])
    j = np.arange(ng, dtype=np.float64)
    table = np.rint((1.0 + A * np.cos(2 * np.pi * j / ng)) * 2.0 ** 24).astype(np.int64)
    ax = np.arange(ng, dtype=np.int64)
    idx = ((m_[0] * ax[:, None, None] + m_[1] * ax[None, :, None]) % ng + m_[2] * ax[None, None, :]) % ng
    rho = table[idx]
    del idx
    with mk(hg, ng, 32) as m:
        if plan:
            m.set_tunable("fft_plan", plan)
        if xs >= 0:
            m.set_tunable("fft_xsolve", xs)
            m.set_tunable("fft_cols", xs)
        m.set_density(rho)
        del rho
        m.solve_gradient()
        F = m.get_force()
    rng = np.random.default_rng(ng)
    s = rng.integers(0, ng ** 3, 1 << 21)
    x, y, z = s // (ng * ng), (s // ng) % ng, s % ng
    k = 2 * np.pi * m_ / ng
    phase = 2 * np.pi * ((m_[0] * x + m_[1] * y + m_[2] * z) % ng) / ng
    lam = 2 * np.sum(np.cos(k) - 1)
    err2 = ref2 = 0.0
    for d in range(3):
        expect = A * np.sin(k[d]) * np.sin(phase) / lam
        err2 += float(np.sum((F[s, d].astype(np.float64) - expect) ** 2))
        ref2 += float(np.sum(expect ** 2))
    assert np.sqrt(err2 / ref2) < 2e-5, np.sqrt(err2 / ref2)


def test_fft_plan_tunable(hg):
    """Radix orders: two at 1024 (the autotuner times both and keeps one), one at 512
    (selecting a second is refused)."""
    case("fft plan tunable")
    with mk(hg, 1024, 32) as m:
        assert m.get_tunable("fft_plans") == 2
        m.set_tunable("fft_plan", 1)
        assert m.get_tunable("fft_plan") == 1
        m.set_tunable("fft_plan", 0)
        assert m.get_tunable("fft_plan") == 0
        m.autotune(1)
        assert m.get_tunable("fft_plan") in (0, 1)
    with mk(hg, 512, 32) as m:
        assert m.get_tunable("fft_plans") == 1
        with pytest.raises(hg.HgpmError):
            m.set_tunable("fft_plan", 1)
        assert m.get_tunable("fft_plan") == 0


@pytest.mark.parametrize("ng", [32, 96, 128])
def test_fft_variants_agree(hg, orc, ng):
    """Every column-pass / x-pass variant the autotuner may pick gives the same physics."""
    case(f"fft variants ng={ng}")
    rng = np.random.default_rng(ng + 21)
    with mk(hg, ng) as m:
        nv = m.get_tunable("fft_variants")
        mask = m.get_tunable("fft_variant_mask")
        assert nv == 6 and mask & 0b100
        cell, off = rand_split(rng, m.N, ng, clustered=True)
        m.set_particles_split(cell, off, np.zeros((m.N, 3), np.float32))
        m.cic()
        q = m.get_density()
        Fo = orc.solve_gradient(ng, q * 2.0 ** -24)
        x = (rng.standard_normal((ng,) * 3) + 1j * rng.standard_normal((ng,) * 3)).astype(np.complex64)
        for v in range(nv):
            if not (mask >> v) & 1:
                with pytest.raises(hg.HgpmError):
                    m.set_tunable("fft_cols", v)
                continue
            m.set_tunable("fft_cols", v)
            m.set_tunable("fft_xsolve", v)
            m.set_density(q)
            m.solve_gradient()
            F = m.get_force().astype(np.float64)
            err = np.sqrt(((F - Fo) ** 2).mean()) / np.sqrt((Fo ** 2).mean())
            assert err < 2e-5, (v, err)
            assert rel_l2(m.fft(x, -1), np.fft.fftn(x.astype(np.complex128))) < 2e-6, v
        res = m.autotune(2)
        assert (mask >> res["fft_cols"]) & 1 and (mask >> res["fft_xsolve"]) & 1
        assert all((t is None) == (not (mask >> v) & 1) for v, t in enumerate(res["cols_ms"]))


@pytest.mark.parametrize("ng", [32, 64])
def test_solve_fp64_path(hg, orc, ng):
    case(f"solve fp64 ng={ng}")
    rng = np.random.default_rng(ng + 5)
    rho = 1 + 0.3 * rng.standard_normal((ng,) * 3)
    with mk(hg, ng, prec=1) as m:
        m.set_density_f64(rho)
        m.solve_gradient()
        F = m.get_force()
    Fo = orc.solve_gradient(ng, rho)
    assert rel_l2(F, Fo) < 1e-12


# --------------------------------------------------------- kick / drift ----
@pytest.mark.parametrize("ng,np_,dist", [(32, 32, "random"), (64, 64, "random"), (64, 64, "lattice"),
                                         (128, 64, "lattice"), (64, 64, "sparse"), (512, 32, "lattice"),
                                         (96, 96, "random"), (96, 32, "lattice"), (192, 64, "sparse")])
def test_kick_drift_bit_exact(hg, orc, ng, np_, dist):
    """Kick and drift against the oracle, bit for bit.  "sparse": displacements of 1.5
    cells fill each gather tile's box thinly, so the gather stages only the rows its
    particles read (row-masked slabs).  (512, 32): particles 16 cells apart, so a tile's
    box spans hundreds of cells and is staged in many slabs."""
    case(f"kick/drift ng={ng} np={np_} {dist}")
    rng = np.random.default_rng(ng + 11)
    with mk(hg, ng, np_) as m:
        if dist == "random":
            cell, off = rand_split(rng, m.N, ng)
        elif dist == "sparse":
            cell, off = lattice_split(rng, np_, ng, 1.5)
        else:
            cell, off = lattice_split(rng, np_, ng, 2.5, outliers=2)
        vel = rng.normal(0, 3, (m.N, 3)).astype(np.float32)
        m.set_particles_split(cell, off, vel)
        m.cic()
        m.solve_gradient()
        F = m.get_force()
        dT, fs, pf = 0.0123, 1.87, 0.75
        vref = orc.kick_split(ng, cell, off, vel, F, dT, fs)
        # kick alone, with every gather workgroup shape, staging from the deposit's cell
        # boxes and (after a zero drift that invalidates them) from its own reduction
        for gt in (1, 0):
            for boxed in (True, False):
                m.set_tunable("gather_tile", gt)
                m.set_particles_split(cell, off, vel)
                m.cic()
                m.solve_gradient()
                if not boxed:
                    m.update_positions(0.0)
                m.update_velocities(dT, fs)
                c1, o1, v1 = m.get_particles_split()
                assert np.array_equal(v1, vref), (gt, boxed)
                assert np.array_equal(c1, cell) and np.array_equal(o1, off)
        # fused kick + drift from the same start
        m.set_particles_split(cell, off, vel)
        m.cic()
        m.solve_gradient()
        m.kick_drift(dT, fs, pf)
        c2, o2, v2 = m.get_particles_split()
        cref, oref = orc.drift_split(ng, cell, off, vref, pf)
        assert np.array_equal(v2, vref)
        assert np.array_equal(c2, cref) and np.array_equal(o2, oref)
        # standalone drift
        m.update_positions(-0.4)
        c3, o3, _ = m.get_particles_split()
        c3r, o3r = orc.drift_split(ng, cref, oref, vref, -0.4)
        assert np.array_equal(c3, c3r) and np.array_equal(o3, o3r)


@pytest.mark.parametrize("ng,np_,dist", [(64, 64, "random"), (96, 32, "random"), (1024, 64, "lattice")])
def test_fused_drift_large_moves(hg, orc, ng, np_, dist):
    """The drift inside the gather with moves of up to three boxes, bit-exact against
    the oracle: moves reduced modulo ng, moves of exactly ng/2, ng and 2 ng, and at
    ng = 1024 moves past 512 cells, where the power-of-two path's 32-bit sums wrap
    (ng = 96 takes the general path).  A zero kick leaves the chosen velocities exact."""
    case(f"fused drift large moves ng={ng} np={np_}")
    rng = np.random.default_rng(5 * ng + np_)
    with mk(hg, ng, np_) as m:
        if dist == "random":
            cell, off = rand_split(rng, m.N, ng)
        else:
            cell, off = lattice_split(rng, np_, ng, 1.5)
        vel = (rng.uniform(-3.0, 3.0, (m.N, 3)) * ng).astype(np.float32)
        special = np.array([0.0, ng / 2, -ng / 2, ng - 2.0 ** -10, -(ng - 2.0 ** -10), ng, -ng, 2 * ng,
                            -3 * ng, 511.99997, 512.0, -512.0, -512.00006, 1023.99994, 2.0 ** -30],
                           np.float32)
        vel.reshape(-1)[:special.size] = special
        vel.reshape(-1)[special.size:2 * special.size] = special[::-1]
        m.set_particles_split(cell, off, vel)
        m.cic()
        m.solve_gradient()
        m.kick_drift(0.0, 1.0, 1.0)  # v + 0 * S = v for every finite force
        c2, o2, v2 = m.get_particles_split()
        cref, oref = orc.drift_split(ng, cell, off, vel, 1.0)
        assert np.array_equal(v2, vel)
        assert np.array_equal(c2, cref) and np.array_equal(o2, oref)


def test_fp64_kick_drift(hg, orc):
    case("kick/drift fp64 ng=32")
    ng = 32
    rng = np.random.default_rng(2)
    pos = rng.random((ng ** 3, 3)) * ng
    vel = rng.normal(0, 2, (ng ** 3, 3))
    with mk(hg, ng, prec=1) as m:
        m.set_particles_f64(pos, vel)
        m.cic()
        rho = m.get_density_f64()
        assert np.abs(rho - orc.cic_f64(ng, pos)).max() < 1e-10
        m.solve_gradient()
        F = m.get_force()
        m.update_velocities(0.02, 1.3)
        _, v = m.get_particles_f64()
        np.testing.assert_allclose(v, orc.kick_f64(ng, pos, vel, F, 0.02, 1.3), rtol=0, atol=1e-12)
        m.update_positions(0.5)
        p2, _ = m.get_particles_f64()
        pf = float(np.float32(0.5))
        np.testing.assert_allclose(p2, orc.drift_f64(ng, pos, v, pf), rtol=0, atol=1e-11)


# ------------------------------------------------------------------ P(k) ----
@pytest.mark.parametrize("ng,folds", [(32, 0), (64, 0), (64, 2)])
def test_power_spectrum(hg, orc, ng, folds):
    case(f"power ng={ng} folds={folds}")
    rng = np.random.default_rng(ng + folds)
    nb = 24
    with mk(hg, ng, pk_bins=nb, pk_folds=folds) as m:
        cell, off = rand_split(rng, m.N, ng, clustered=True)
        m.set_particles_split(cell, off, np.zeros((m.N, 3), np.float32))
        for f in range(folds + 1):
            k, pk, cnt = m.power_spectrum(f)
            c, o = cell, off
            for _ in range(f):
                c, o = orc.fold_split(ng, c, o)
            q = orc.cic_split(ng, c, o)
            ko, pko, cnto = orc.power_grid(ng, 2.0 * ng, folds, f, nb, q * 2.0 ** -24)
            assert np.array_equal(cnt, cnto)
            np.testing.assert_array_equal(k, ko)
            nz = cnto > 0
            assert np.abs(pk[nz] / pko[nz] - 1).max() < 1e-4
        # deterministic FP64 sums
        k2, pk2, cnt2 = m.power_spectrum(folds)
        assert np.array_equal(pk2, pk) and np.array_equal(cnt2, cnt)


def test_power_spectrum_fp64(hg, orc):
    case("power fp64 ng=32")
    ng, nb = 32, 16
    rng = np.random.default_rng(8)
    pos = rng.random((ng ** 3, 3)) * ng
    with mk(hg, ng, prec=1, pk_bins=nb, pk_folds=1) as m:
        m.set_particles_f64(pos, np.zeros_like(pos))
        for f in (0, 1):
            k, pk, cnt = m.power_spectrum(f)
            pp = pos.copy()
            for _ in range(f):
                pp = orc.fold_f64(ng, pp)
            ko, pko, cnto = orc.power_grid(ng, 64.0, 1, f, nb, orc.cic_f64(ng, pp))
            assert np.array_equal(cnt, cnto)
            nz = cnto > 0
            assert np.abs(pk[nz] / pko[nz] - 1).max() < 1e-10


# -------------------------------------------------------------------- IC ----
def ipk_for(hg, p, tmp):
    fn = os.path.join(tmp, "eh.ipk")
    hg.write_ipk_eh(p, fn, 50.0, 50001)
    return hg.read_ipk(fn)


@pytest.mark.parametrize("prec", [0, 1])
def test_initial_conditions(hg, orc, tmp_path, prec):
    case(f"ic ng=64 prec={prec}")
    p = hg.make_params(64, 64, 128.0, 50.0, 0.0, 100, seed=21082001, precision=prec)
    hdr, vals = ipk_for(hg, p, str(tmp_path))
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    with hg.MemoryManager(p) as m:
        D, dD = hg.GenerateDisplacementIC(m, p, ts, hdr, vals)
        if prec == 0:
            c, o, v = m.get_particles_split()
            x = c + o.astype(np.float64)
        else:
            x, v = m.get_particles_f64()
    xo, vo = orc.ic_f64(64, 64, 128.0, 21082001, vals, hdr[3], D, dD, p.z_ini, ts.fscal, ts.deltaT)
    d = (x - xo + 32) % 64 - 32
    disp = (xo - np.indices((64,) * 3).reshape(3, -1).T + 32) % 64 - 32
    # FP32: float noise grid, FP32 transforms and float k (reference rounding);
    # FP64: the oracle's recipe in FP64
    tol_x, tol_v = (1e-4, 1e-4) if prec == 0 else (1e-9, 1e-9)
    print(f"[result] ic max|dx| {np.abs(d).max():.3e} max rel dv {np.abs(v - vo).max() / np.abs(vo).max():.3e}")
    assert np.abs(d).max() < tol_x
    assert np.abs(v - vo).max() < tol_v * np.abs(vo).max()
    assert np.sqrt((disp ** 2).mean()) > 1e-2  # non-trivial displacements


# ------------------------------------------------------------ step path ----
@pytest.mark.parametrize("ng,np_", [(64, 64), (128, 64)])
def test_step_fused_deposit_matches_separate_calls(hg, orc, ng, np_):
    """hgpm_step (merged drifts, an output step with half drifts and P(k) in between)
    must leave particles and density bit-identical to the step done by separate calls.
    The step's owner pass hands the solve an FP32 density when no tile is deferred; a
    second pair of contexts with a tiny deposit arena defers tiles and takes the int64
    route inside the step."""
    case(f"fused step ng={ng} np={np_}")
    rng = np.random.default_rng(ng + np_)
    n_steps = 6
    p = hg.make_params(ng, np_, 2.0 * ng, 50.0, 0.0, 100)
    out_after = np.zeros(n_steps, np.uint8)
    out_after[2] = 1
    sched = hg.schedule(p, n_steps, merge=True, output_after=out_after)
    cell, off = lattice_split(rng, np_, ng, 1.5, outliers=2)
    vel = rng.normal(0, 2.0, (np_ ** 3, 3)).astype(np.float32)
    for arena in (None, 1):
        with hg.MemoryManager(p) as m1, hg.MemoryManager(p) as m2:
            if arena is not None:
                m1.set_tunable("deposit_arena", arena)
                m2.set_tunable("deposit_arena", arena)
            m1.set_particles_split(cell, off, vel)
            m2.set_particles_split(cell, off, vel)
            for s, co in enumerate(sched):
                m1.step(co)
                if not co.skip_drift0:
                    m2.update_positions(co.drift0)
                m2.cic()
                m2.solve_gradient()
                m2.kick_drift(co.kick_dt, co.kick_fscal, co.drift1)
                if out_after[s]:
                    k1, pk1, c1 = m1.power_spectrum(0)
                    k2, pk2, c2 = m2.power_spectrum(0)
                    assert np.array_equal(pk1, pk2) and np.array_equal(c1, c2)
                a, b = m1.get_particles_split(), m2.get_particles_split()
                for x, y in zip(a, b):
                    assert np.array_equal(x, y), s
            if arena is not None:
                assert m1.get_tunable("deposit_deferred") > 0
            m1.cic()
            m2.cic()
            assert np.array_equal(m1.get_density(), m2.get_density())


def test_step_results_independent_of_particle_order(hg):
    """The same particles in lattice order (compact tiles, arena route) and in a random
    order (every tile spans the box: atomic route): bit-identical particles and P(k)
    after three steps, up to the permutation.  Integer deposit sums are order-free and
    the gather treats each particle alone."""
    case("step, particle order")
    ng = np_ = 64
    rng = np.random.default_rng(64)
    p = hg.make_params(ng, np_, 2.0 * ng, 50.0, 0.0, 100)
    cell, off = lattice_split(rng, np_, ng, 1.5)
    vel = rng.normal(0, 0.01, (np_ ** 3, 3)).astype(np.float32)  # the tiles stay compact
    perm = rng.permutation(np_ ** 3)
    with hg.MemoryManager(p) as a, hg.MemoryManager(p) as b:
        a.set_particles_split(cell, off, vel)
        b.set_particles_split(cell[perm], off[perm], vel[perm])
        for co in hg.schedule(p, 3, merge=True):
            a.step(co)
            b.step(co)
        assert a.get_tunable("deposit_deferred") == 0
        assert b.get_tunable("deposit_deferred") > 0
        for x, y in zip(a.get_particles_split(), b.get_particles_split()):
            assert np.array_equal(x[perm], y)
        assert np.array_equal(a.power_spectrum(0)[1], b.power_spectrum(0)[1])


# -------------------------------------------------------------- full run ----
def test_full_run_c1_fp32_and_fp64_vs_oracle(hg, orc, tmp_path):
    """64^3 particles, 64^3 grid, box 128 Mpc/h, z=200 -> 0 (the reference's CPU-runnable
    configuration) in 120 a-uniform steps -- with 60 steps the reference's initial
    reverseHalfStep makes a < 0 -- : GPU FP32 engine and GPU FP64 plain path against the
    FP64 oracle from identical initial conditions."""
    case("full run 64^3 x 120")
    ng, n = 64, 120
    p = hg.make_params(ng, ng, 128.0, 200.0, 0.0, n, seed=21082001, pk_bins=32)
    hdr, vals = ipk_for(hg, p, str(tmp_path))
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    a = 1 / (1 + p.z_ini)
    D, _ = hg.growth(p, a)
    _, dD = hg.growth(p, a - ts.deltaT / 2)
    pos, vel = orc.ic_f64(ng, ng, 128.0, 21082001, vals, hdr[3], D, dD, p.z_ini, ts.fscal, ts.deltaT)
    co = hg.schedule_array(p, n, merge=False)
    po, vo, _ = orc.run_f64(ng, ng, co, pos, vel)
    q_o = orc.cic_f64(ng, po)
    k_o, pk_o, cnt_o = orc.power_grid(ng, 128.0, 0, 0, 32, q_o)
    # FP32 engine, merged drifts
    sched = hg.schedule(p, n, merge=True)
    cell, off = orc.split_positions(ng, pos)
    with hg.MemoryManager(p) as m:
        m.set_particles_split(cell, off, vel.astype(np.float32))
        for c in sched:
            m.step(c)
        c32, o32, v32 = m.get_particles_split()
        k, pk, cnt = m.power_spectrum(0)
    x32 = c32 + o32.astype(np.float64)
    d = (x32 - po + ng / 2) % ng - ng / 2
    rms = float(np.sqrt((d ** 2).mean()))
    kny = np.pi * ng / 128.0
    sel = (cnt_o > 0) & (k_o < kny / 2)
    perr = float(np.abs(pk[sel] / pk_o[sel] - 1).max())
    print(f"[result] fp32 rms displacement {rms:.3e} cells, max rel P(k) err {perr:.3e}", flush=True)
    assert np.array_equal(cnt, cnt_o)
    assert rms < 2e-3 and perr < 1e-3
    # FP64 plain path
    p64 = hg.make_params(ng, ng, 128.0, 200.0, 0.0, n, seed=21082001, pk_bins=32, precision=1)
    with hg.MemoryManager(p64) as m:
        m.set_particles_f64(pos, vel)
        for c in hg.schedule(p64, n, merge=False):
            m.step(c)
        x64, v64 = m.get_particles_f64()
    d = (x64 - po + ng / 2) % ng - ng / 2
    assert np.sqrt((d ** 2).mean()) < 1e-9


def test_driver_end_to_end(hg, tmp_path):
    """haccgpm <params> writes the reference's outputs: <prefix>.params, P(k) text files
    (k,P,count per line) for ini/steps/fin, raw float3 snapshots."""
    case("driver 32^3 x 10")
    import subprocess
    p = hg.make_params(32, 32, 64.0, 50.0, 0.0, 100)
    ipk = tmp_path / "eh.ipk"
    hg.write_ipk_eh(p, str(ipk), 20.0, 20001)
    prm = tmp_path / "run.params"
    prm.write_text("\n".join([
        "OMEGA_CDM 0.261", "DEUT 0.0226576", "OMEGA_NU 0.0", "HUBBLE 0.68", "SS8 0.81", "NS 0.96",
        "W_DE -1.0", "WA_DE 0.0", "T_CMB 2.726", "N_EFF_MASSLESS 3.04", "N_EFF_MASSIVE 0.0",
        f"OUTPUT_BASE_NAME {tmp_path}/out", f"IPK {ipk}", "Z_IN 50", "Z_FIN 0", "N_STEPS 100", "LAST_STEP 10",
        "PK_FOLDS 1", "PK_BINS 32", "PK_DUMP 0,4", "PARTICLE_DUMP 4", "DUMP_INIT", "DUMP_FINAL",
        "NG 32", "NP 32", "RL 64", "SEED 5", "BLOCK_SIZE 64", ""]))
    r = subprocess.run([hg.DRIVER_PATH, str(prm)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    for name in ["out.params", "out.pk.ini.fold.0", "out.pk.ini.fold.1", "out.pk.0.fold.0", "out.pk.4.fold.1",
                 "out.pk.fin.fold.0", "out.particles.ini", "out.particles.4", "out.particles.fin"]:
        assert (tmp_path / name).exists(), name
    rows = [l.split(",") for l in (tmp_path / "out.pk.fin.fold.0").read_text().splitlines()]
    assert len(rows) > 5 and all(len(r) == 3 for r in rows)
    ks = np.array([float(r[0]) for r in rows])
    assert np.all(np.diff(ks) > 0) and all(int(r[2]) > 0 for r in rows)
    snap = np.fromfile(tmp_path / "out.particles.fin", np.float32)
    assert snap.size == 6 * 32 ** 3
    pos = snap[: 3 * 32 ** 3]
    assert pos.min() >= 0 and pos.max() < 32


def test_driver_analysis_hook(hg, tmp_path):
    """DO_ANALYSIS with ANALYSIS_STEPS 2,5 (params.cpp:195-214, driver_pm.cpp:153): at
    those steps and only there the driver writes the 7-float-per-particle array of
    serial::PyAnalysis (py_analysis.cu:3-26) and runs ANALYSIS_PY on it with the step,
    a and z after the step.  The step-5 array holds the index and the same float3
    positions and velocities as the step-5 snapshot."""
    case("driver analysis hook 32^3 x 7")
    import json
    import subprocess
    p = hg.make_params(32, 32, 64.0, 50.0, 0.0, 100)
    ipk = tmp_path / "eh.ipk"
    hg.write_ipk_eh(p, str(ipk), 20.0, 20001)
    adir = tmp_path / "ana"
    adir.mkdir()
    script = tmp_path / "ana.py"
    script.write_text(
        "import json, sys\nimport numpy as np\n"
        "fn, n, w, step, a, z, d = sys.argv[1:8]\n"
        "x = np.fromfile(fn, np.float32).reshape(int(n), int(w))\n"
        "json.dump({'n': int(n), 'w': int(w), 'step': int(step), 'a': float(a), 'z': float(z), 'dir': d,\n"
        "           'index_ok': bool((x[:, 0] == np.arange(int(n))).all()), 'pmin': float(x[:, 1:4].min()),\n"
        "           'pmax': float(x[:, 1:4].max())}, open(d + '/seen.' + step + '.json', 'w'))\n")
    prm = tmp_path / "run.params"
    prm.write_text("\n".join([
        "OMEGA_CDM 0.261", "DEUT 0.0226576", "OMEGA_NU 0.0", "HUBBLE 0.68", "SS8 0.81", "NS 0.96",
        "W_DE -1.0", "WA_DE 0.0", "T_CMB 2.726", "N_EFF_MASSLESS 3.04", "N_EFF_MASSIVE 0.0",
        f"OUTPUT_BASE_NAME {tmp_path}/out", f"IPK {ipk}", "Z_IN 50", "Z_FIN 0", "N_STEPS 100", "LAST_STEP 7",
        "PK_BINS 32", "PARTICLE_DUMP 5", "ANALYSIS_STEPS 2,5", f"ANALYSIS_DIR {adir}", f"ANALYSIS_PY {script}",
        "DO_ANALYSIS", "NG 32", "NP 32", "RL 64", "SEED 5", ""]))
    env = dict(os.environ, HGPM_PYTHON=sys.executable)
    r = subprocess.run([hg.DRIVER_PATH, str(prm)], capture_output=True, text=True, timeout=120, env=env)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    assert sorted(f.name for f in adir.iterdir()) == ["out.analysis.2", "out.analysis.5", "seen.2.json",
                                                      "seen.5.json"]
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    ts.advanceHalfStep()
    for step in range(6):
        ts.advanceHalfStep()
        ts.advanceHalfStep()
        if step in (2, 5):
            seen = json.load(open(adir / f"seen.{step}.json"))
            assert seen["n"] == 32 ** 3 and seen["w"] == 7 and seen["step"] == step and seen["index_ok"]
            assert seen["a"] == ts.aa and seen["z"] == ts.z and seen["dir"] == str(adir)
            assert 0 <= seen["pmin"] and seen["pmax"] < 32
    ana = np.fromfile(adir / "out.analysis.5", np.float32).reshape(-1, 7)
    snap = np.fromfile(tmp_path / "out.particles.5", np.float32)
    n = 32 ** 3
    assert np.array_equal(ana[:, 1:4], snap[:3 * n].reshape(n, 3))
    assert np.array_equal(ana[:, 4:7], snap[3 * n:].reshape(n, 3))


# -------------------------------------------------------------- ensemble ----
def _ic_inputs(hg, p):
    import tempfile
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "eh.ipk")
        hg.write_ipk_eh(p, fn, 100.0, 20001)
        hdr, vals = hg.read_ipk(fn)
    a = 1.0 / (1.0 + p.z_ini)
    delta, _ = hg.growth(p, a)
    _, dot_delta = hg.growth(p, a - ts.deltaT / 2.0)
    return hdr, vals, delta, dot_delta, ts


def test_ensemble_matches_single_boxes(hg):
    """Boxes stepped as a batch (one launch per kernel for the batch) and as separate
    single-box contexts: bit-identical particles and P(k)."""
    case("ensemble 2 x 32^3 (batched) + 64^3")
    ps = [hg.make_params(32, 32, 64.0, 30.0, 0.0, 40, seed=11, ss8=0.80),
          hg.make_params(32, 32, 64.0, 30.0, 0.0, 40, seed=12, ss8=0.75),
          hg.make_params(64, 64, 128.0, 30.0, 0.0, 40, seed=13, ss8=0.85)]
    n_steps = 5
    with hg.Ensemble(ps) as e:
        nb, of = e.batches()
        assert nb == 2 and of[0] == of[1] != of[2]
        singles = [hg.MemoryManager(p) for p in ps]
        try:
            scheds = []
            for b, p in enumerate(ps):
                hdr, vals, delta, dot_delta, ts = _ic_inputs(hg, p)
                e.generate_ic(b, hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT)
                singles[b].generate_ic(hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT)
                scheds.append(hg.schedule(p, n_steps, merge=True))
            for s in range(n_steps):
                e.step([scheds[b][s] for b in range(len(ps))])
                for b in range(len(ps)):
                    singles[b].step(scheds[b][s])
            for b in range(len(ps)):
                pe, ve = e.get_particles(b)
                p1, v1 = singles[b].get_particles()
                assert np.array_equal(pe, p1) and np.array_equal(ve, v1), b
                assert np.array_equal(e.power_spectrum(b)[1], singles[b].power_spectrum(0)[1]), b
            # boxes in one batch must share their step coefficients
            bad = [scheds[b][0] for b in range(len(ps))]
            c = hg.StepCoeffs()
            c.drift0, c.kick_dt, c.kick_fscal, c.drift1, c.skip_drift0 = 0.1, 0.2, 0.3, 0.4, 0
            bad[1] = c
            with pytest.raises(hg.HgpmError):
                e.step(bad)
        finally:
            for m in singles:
                m.close()



def test_ensemble_refuses_boxes_without_particles(hg):
    """A batch shares its buffers, but readiness is per box: with one box of a batch
    generated and the other not, stepping, running and reading the unset box are
    refused with HGPM_ERR_STATE (nothing is launched on its uninitialised memory)."""
    case("ensemble 2 x 32^3, one box without particles")
    ps = [hg.make_params(32, 32, 64.0, 30.0, 0.0, 40, seed=21 + b) for b in range(2)]
    with hg.Ensemble(ps) as e:
        assert e.batches()[0] == 1
        hdr, vals, delta, dot_delta, ts = _ic_inputs(hg, ps[0])
        e.generate_ic(0, hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT)
        co = [hg.schedule(p, 2, merge=True)[0] for p in ps]
        for call in (lambda: e.step(co), lambda: e.run([co]), lambda: e.get_particles(1)):
            with pytest.raises(hg.HgpmError) as err:
                call()
            assert err.value.code == hg.ERR_STATE
        e.get_particles(0)  # the generated box reads back
        e.generate_ic(1, hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT)
        e.step(co)
        e.synchronize()

def test_ensemble_graph_run_matches_steps(hg):
    """Steps captured into one HIP graph per run (hgpm_ensemble_run) against the same
    steps launched one by one: bit-identical particles and P(k) after each of three
    runs, with eager per-box work (P(k), readback) between the runs.  Two batches on
    two streams, and a third ensemble stored in waves."""
    case("ensemble graph run: 2 x 32^3 + 64^3, and 4 x 64^3 in waves")
    ps = [hg.make_params(32, 32, 64.0, 30.0, 0.0, 40, seed=31, ss8=0.80),
          hg.make_params(32, 32, 64.0, 30.0, 0.0, 40, seed=32, ss8=0.72),
          hg.make_params(64, 64, 128.0, 30.0, 0.0, 40, seed=33, ss8=0.88)]
    qs = [hg.make_params(64, 64, 128.0, 30.0, 0.0, 40, seed=40 + b, ss8=0.7 + 0.05 * b) for b in range(4)]
    n_steps, seg = 6, 2
    with hg.Ensemble(ps) as eager, hg.Ensemble(ps) as graph, \
            hg.Ensemble(qs) as q_eager, hg.Ensemble(qs, device_budget=120 << 20) as q_graph:
        assert eager.batches()[0] == 2
        assert int(q_graph.waves()[0]) == 2
        pairs = [(eager, graph, ps), (q_eager, q_graph, qs)]
        scheds = {}
        for a, g, pl in pairs:
            sc = []
            for b, p in enumerate(pl):
                hdr, vals, delta, dot_delta, ts = _ic_inputs(hg, p)
                a.generate_ic(b, hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT)
                g.generate_ic(b, hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT)
                sc.append(hg.schedule(p, n_steps, merge=True))
            scheds[id(a)] = sc
        for s0 in range(0, n_steps, seg):
            for a, g, pl in pairs:
                sc = scheds[id(a)]
                for s in range(s0, s0 + seg):
                    a.step([sc[b][s] for b in range(len(pl))])
                g.run([[sc[b][s] for b in range(len(pl))] for s in range(s0, s0 + seg)])
            for a, g, pl in pairs:
                for b in range(len(pl)):
                    p1, v1 = a.get_particles(b)
                    p2, v2 = g.get_particles(b)
                    assert np.array_equal(p1, p2) and np.array_equal(v1, v2), (s0, b)
                    assert np.array_equal(a.power_spectrum(b)[1], g.power_spectrum(b)[1]), (s0, b)
        graph.synchronize()
        bad = [scheds[id(eager)][b][0] for b in range(len(ps))]
        c = hg.StepCoeffs()
        c.drift0, c.kick_dt, c.kick_fscal, c.drift1, c.skip_drift0 = 0.1, 0.2, 0.3, 0.4, 0
        bad[1] = c
        with pytest.raises(hg.HgpmError):
            graph.run([bad])


def test_ensemble_waves_match_resident(hg):
    """An ensemble squeezed into a small device budget keeps every box's particles but
    grid buffers for a wave of boxes only, and steps the waves in turn: bit-identical
    to the same ensemble fully resident.  A budget below one box per wave is refused."""
    case("ensemble 6 x 64^3 in waves")
    ps = [hg.make_params(64, 64, 128.0, 30.0, 0.0, 40, seed=20 + b, ss8=0.75 + 0.02 * b) for b in range(6)]
    n_steps = 4
    with hg.Ensemble(ps) as full, hg.Ensemble(ps, device_budget=150 << 20) as waved:
        assert full.waves()[0] == 6
        w = int(waved.waves()[0])
        assert 1 <= w < 6, w
        assert waved.device_bytes() < full.device_bytes()
        scheds = []
        for b, p in enumerate(ps):
            hdr, vals, delta, dot_delta, ts = _ic_inputs(hg, p)
            full.generate_ic(b, hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT)
            waved.generate_ic(b, hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT)
            scheds.append(hg.schedule(p, n_steps, merge=True))
        for s in range(n_steps):
            full.step([scheds[b][s] for b in range(len(ps))])
            waved.step([scheds[b][s] for b in range(len(ps))])
        for b in range(len(ps)):
            p1, v1 = full.get_particles(b)
            p2, v2 = waved.get_particles(b)
            assert np.array_equal(p1, p2) and np.array_equal(v1, v2), b
            assert np.array_equal(full.power_spectrum(b)[1], waved.power_spectrum(b)[1]), b
    with pytest.raises(hg.HgpmError):
        hg.Ensemble(ps, device_budget=16 << 20)