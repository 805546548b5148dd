"""Pins of the CPU oracle (oracle/hgpm_oracle.c) against known answers, closed forms,
invariants and brute force -- never against a retyped copy of its own formulas."""
import itertools
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, rand_split


# ---------------------------------------------------------------- RNG ----
def test_philox_known_answers(orc):
    # Random123 kat_vectors for philox4x32-10
    assert [hex(x) for x in orc.philox([0, 0, 0, 0], [0, 0])] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(x) for x in orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == [
        "0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(x) for x in orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0])] == [
        "0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_det_log_cos_match_libm(orc):
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.random(20000), [1.0, 0.5, 2.0 ** -53, 0.7071067811865476, 1 - 2 ** -53]])
    for x in xs:
        assert abs(orc.L().orc_det_log(float(x)) - math.log(x)) <= 4e-15 * max(1.0, abs(math.log(x)))
    for u in np.concatenate([rng.random(20000), [0.0, 0.25, 0.5, 0.75, 0.125]]):
        assert abs(orc.L().orc_det_cos2pi(float(u)) - math.cos(2 * math.pi * u)) <= 2e-15
    assert orc.L().orc_det_cos2pi(0.0) == 1.0 and orc.L().orc_det_log(1.0) == 0.0


def test_white_noise_moments(orc):
    z = orc.white_noise(21082001, 1 << 20)
    assert abs(z.mean()) < 5 / math.sqrt(z.size)
    assert abs(z.var() - 1) < 5 * math.sqrt(2 / z.size)
    # tail fractions of N(0,1)
    assert abs((np.abs(z) > 2).mean() - 0.0455) < 0.002
    # seeds decorrelate, reproducible
    z2 = orc.white_noise(21082002, 1 << 16)
    assert abs(np.corrcoef(z[: 1 << 16], z2)[0, 1]) < 0.02
    assert np.array_equal(orc.white_noise(21082001, 1000), z[:1000])


# ---------------------------------------------------------------- FFT ----
def test_fft_numpy_golden(orc):
    g = np.load(os.path.join(GOLDEN, "fft_golden.npz"))
    for n in (8, 12, 16):
        X = orc.fft3d(g[f"x{n}"], -1)
        np.testing.assert_allclose(X, g[f"X{n}"], rtol=0, atol=1e-11 * np.abs(g[f"X{n}"]).max())


def test_fft_brute_force_and_identities(orc):
    rng = np.random.default_rng(3)
    n = 6
    x = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    idx = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(idx, idx) / n)
    brute = np.einsum("ai,bj,ck,ijk->abc", W, W, W, x)
    np.testing.assert_allclose(orc.fft3d(x, -1), brute, atol=1e-11)
    X = orc.fft3d(x, -1)
    assert abs(np.sum(np.abs(X) ** 2) - n ** 3 * np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(X) ** 2)
    np.testing.assert_allclose(orc.fft3d(X, +1), n ** 3 * x, atol=1e-10)
    # delta -> constant; plane wave -> spike
    d = np.zeros((n, n, n), complex); d[0, 0, 0] = 1
    np.testing.assert_allclose(orc.fft3d(d, -1), np.ones((n, n, n)), atol=1e-14)
    m = (1, 2, 5)
    I = np.indices((n, n, n))
    pw = np.exp(2j * np.pi * (m[0] * I[0] + m[1] * I[1] + m[2] * I[2]) / n)
    S = orc.fft3d(pw, -1)
    assert abs(S[m] - n ** 3) < 1e-9 and np.abs(S).sum() - abs(S[m]) < 1e-8


# ------------------------------------------------------------ deposit ----
def test_deposit_closed_forms(orc):
    ng = 8
    # on-lattice particles: one cell each, exactly 2^24
    cell = np.array([[0, 0, 0], [7, 7, 7], [3, 4, 5]], np.int32)
    off = np.zeros((3, 3), np.float32)
    rho = orc.cic_split(ng, cell, off)
    assert rho.sum() == 3 << 24 and rho[0, 0, 0] == 1 << 24 and rho[7, 7, 7] == 1 << 24 and rho[3, 4, 5] == 1 << 24
    # particle at the cell centre of (7,7,7): 8 equal weights of 2^21, wrapped periodically
    rho = orc.cic_split(ng, np.array([[7, 7, 7]], np.int32), np.full((1, 3), 0.5, np.float32))
    nz = np.argwhere(rho)
    assert len(nz) == 8 and set(map(tuple, nz)) == set(itertools.product((7, 0), repeat=3))
    assert np.all(rho[rho != 0] == 1 << 21)


def test_deposit_invariants(orc):
    rng = np.random.default_rng(5)
    ng, n = 16, 5000
    cell, off = rand_split(rng, n, ng)
    rho = orc.cic_split(ng, cell, off)
    # each particle deposits 2^24 up to 8 roundings of half a unit
    assert abs(int(rho.sum()) - n * (1 << 24)) <= 4 * n
    # permutation invariance (exact) and additivity (exact)
    perm = rng.permutation(n)
    assert np.array_equal(orc.cic_split(ng, cell[perm], off[perm]), rho)
    a = orc.cic_split(ng, cell[: n // 2], off[: n // 2])
    b = orc.cic_split(ng, cell[n // 2:], off[n // 2:])
    assert np.array_equal(a + b, rho)
    # translation by one cell = periodic roll
    assert np.array_equal(orc.cic_split(ng, (cell + [1, 0, 2]) % ng, off), np.roll(rho, (1, 0, 2), (0, 1, 2)))
    # agreement with the FP64 deposit to the fixed-point rounding
    pos = cell.astype(np.float64) + off
    r64 = orc.cic_f64(ng, pos)
    assert np.abs(rho * 2.0 ** -24 - r64).max() <= (8 * 0.5 * 2.0 ** -24 + 1e-6) * 8


def test_deposit_brute_force_tiny(orc):
    # brute force CIC from the definition: weight = prod_d (1 - |x_d - c_d|) over the 8 nearest cells
    rng = np.random.default_rng(9)
    ng = 4
    cell, off = rand_split(rng, 40, ng)
    x = cell + off.astype(np.float64)
    ref = np.zeros((ng, ng, ng))
    for p in x:
        for cx in range(ng):
            for cy in range(ng):
                for cz in range(ng):
                    w = 1.0
                    for d, c in enumerate((cx, cy, cz)):
                        dist = abs(p[d] - c)
                        dist = min(dist, ng - dist)
                        w *= max(0.0, 1.0 - dist)
                    ref[cx, cy, cz] += w
    rho = orc.cic_split(ng, cell, off) * 2.0 ** -24
    assert np.abs(rho - ref).max() < 1e-6


# --------------------------------------------------------------- solve ----
def test_solve_single_mode_closed_form(orc):
    ng = 16
    m = np.array([1, 2, 3])
    I = np.indices((ng, ng, ng))
    phase = 2 * np.pi * (m[0] * I[0] + m[1] * I[1] + m[2] * I[2]) / ng
    rho = 1.0 + 0.3 * np.cos(phase)
    F = orc.solve_gradient(ng, rho).reshape(ng, ng, ng, 3)
    # rho = 0.3 cos(k.x): phi = 0.3 cos(k.x) / (2 sum(cos k_d - 1)) under the 7-point
    # Laplacian; F_d = -(central difference of phi)_d = 0.3 sin(k_d) sin(k.x) / (2 sum(cos k_d - 1))
    k = 2 * np.pi * m / ng
    lam = 2 * np.sum(np.cos(k) - 1)
    for d in range(3):
        expect = 0.3 * np.sin(k[d]) * np.sin(phase) / lam
        np.testing.assert_allclose(F[..., d], expect, atol=1e-13)


def test_solve_laplacian_and_net_force(orc):
    rng = np.random.default_rng(2)
    ng = 12
    rho = 1 + 0.2 * rng.standard_normal((ng, ng, ng))
    F = orc.solve_gradient(ng, rho).reshape(ng, ng, ng, 3)
    assert np.abs(F.sum(axis=(0, 1, 2))).max() < 1e-10  # G(0) = 0: no net force
    # F = -D phi with D the central difference; the divergence of F with the same stencil
    # equals -(5-point-per-axis Laplacian of phi); check against rho - mean via spectra
    div = sum((np.roll(F[..., d], -1, d) - np.roll(F[..., d], 1, d)) / 2 for d in range(3))
    # in k space: div F^ = -sum sin^2 k_d / (2 sum (cos-1)) rho^ ; compare
    R = np.fft.fftn(rho - rho.mean())
    kk = 2 * np.pi * np.fft.fftfreq(ng)
    KX, KY, KZ = np.meshgrid(kk, kk, kk, indexing="ij")
    lam = 2 * (np.cos(KX) + np.cos(KY) + np.cos(KZ) - 3)
    lam[0, 0, 0] = 1
    ratio = (np.sin(KX) ** 2 + np.sin(KY) ** 2 + np.sin(KZ) ** 2) / lam
    np.testing.assert_allclose(np.fft.fftn(div), R * ratio, atol=1e-9)


def test_solve_is_the_central_difference_of_the_potential(orc):
    """The FP32 engine transforms one field, the potential psi = IFFT(g rho^) with
    g = (0.25/ng^3) / sum_d sin^2(k_d/2), and forms F_d = (psi[c + e_d] - psi[c - e_d]) / 2.
    That is the oracle's spectral gradient -i sin(k_d) G rho^ (G = -g) exactly: the
    identity holds to rounding on a random density and on an odd grid."""
    rng = np.random.default_rng(9)
    for ng in (12, 15):
        rho = 1 + 0.3 * rng.standard_normal((ng, ng, ng))
        F = orc.solve_gradient(ng, rho).reshape(ng, ng, ng, 3)
        k = 2 * np.pi * np.fft.fftfreq(ng)
        s2 = np.sin(k / 2) ** 2
        S = s2[:, None, None] + s2[None, :, None] + s2[None, None, :]
        S[0, 0, 0] = 1.0
        g = 0.25 / S
        g[0, 0, 0] = 0.0
        psi = np.fft.ifftn(g * np.fft.fftn(rho)).real
        for d in range(3):
            fd = (np.roll(psi, -1, d) - np.roll(psi, 1, d)) / 2
            np.testing.assert_allclose(F[..., d], fd, rtol=0, atol=1e-12 * np.abs(F).max())


# ----------------------------------------------------- kick and drift ----
def test_kick_uniform_force_and_linear_field(orc):
    rng = np.random.default_rng(4)
    ng, n = 8, 300
    cell, off = rand_split(rng, n, ng)
    vel = np.zeros((n, 3), np.float32)
    F = np.tile(np.array([0.5, -0.25, 2.0], np.float32), (ng ** 3, 1))
    v = orc.kick_split(ng, cell, off, vel, F, 0.01, 3.0)
    np.testing.assert_allclose(v, np.tile([0.015, -0.0075, 0.06], (n, 1)), rtol=3e-6)
    # CIC interpolates a field linear in x exactly (away from the periodic seam)
    x = np.arange(ng)
    Fx = np.zeros((ng, ng, ng, 3), np.float32)
    Fx[..., 0] = (x[:, None, None] * 0.125).astype(np.float32)
    inside = cell[:, 0] < ng - 1
    v = orc.kick_split(ng, cell, off, vel, Fx.reshape(-1, 3), 1.0, 1.0)
    np.testing.assert_allclose(v[inside, 0], 0.125 * (cell[inside, 0] + off[inside, 0]), rtol=2e-6, atol=1e-6)


def test_drift_wrap_and_merge(orc):
    rng = np.random.default_rng(6)
    ng, n = 16, 2000
    cell, off = rand_split(rng, n, ng)
    vel = rng.normal(0, 40, (n, 3)).astype(np.float32)
    c1, o1 = orc.drift_split(ng, cell, off, vel, 0.37)
    assert c1.min() >= 0 and c1.max() < ng and o1.min() >= 0 and o1.max() < 1
    x0 = cell + off.astype(np.float64)
    x1 = c1 + o1.astype(np.float64)
    dx = (x1 - x0 - vel.astype(np.float64) * np.float32(0.37) + ng / 2) % ng - ng / 2
    assert np.abs(dx).max() < 1e-5
    # two half drifts equal one full drift up to FP32 rounding of the offset
    ca, oa = orc.drift_split(ng, *orc.drift_split(ng, cell, off, vel, 0.25), vel, 0.25)
    cb, ob = orc.drift_split(ng, cell, off, vel, 0.5)
    d = (ca + oa.astype(np.float64)) - (cb + ob.astype(np.float64))
    d = (d + ng / 2) % ng - ng / 2
    assert np.abs(d).max() < 4e-6


# ---------------------------------------------------------------- P(k) ----
def test_power_counts_and_single_mode(orc):
    ng, rl, nb = 16, 32.0, 12
    rng = np.random.default_rng(7)
    rho = 1 + 0.1 * rng.standard_normal((ng, ng, ng))
    k, pk, cnt = orc.power_grid(ng, rl, 0, 0, nb, rho)
    assert cnt.sum() == ng ** 3 - 2  # every mode but k=0 and the (ng/2)^3 corner
    # brute-force counts from the bin layout: |k| in [minK + i*d, minK + (i+1)*d)
    mk, bd, _ = orc.pk_bins(ng, rl, 0, nb)
    kk = 2 * np.pi / rl * np.fft.fftfreq(ng, 1 / ng)
    kk[ng // 2] = -kk[ng // 2]  # get_kmodes maps the Nyquist index to -ng/2
    K = np.sqrt(kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kk[None, None, :] ** 2)
    K[0, 0, 0] = -1
    K[ng // 2, ng // 2, ng // 2] = -1
    b = np.floor((K - mk) / bd)
    b[K < 0] = -9
    b[(b < 0) & (K >= 0)] = 0  # |k| < minK truncates to bin 0
    brute = np.array([(b == i).sum() for i in range(nb)])
    assert np.abs(brute - cnt).sum() <= 8  # float rounding at shell edges
    # single mode: rho = 1 + A cos(2 pi m.x/ng): |rho^(+-m)|^2 = (A ng^3/2)^2
    A, m = 0.4, (2, 0, 1)
    I = np.indices((ng, ng, ng))
    rho1 = 1 + A * np.cos(2 * np.pi * (m[0] * I[0] + m[1] * I[1] + m[2] * I[2]) / ng)
    k, pk, cnt = orc.power_grid(ng, rl, 0, 0, nb, rho1)
    i = int(np.argmax(pk * cnt))
    win = 1.0
    for md in m:
        if md:
            x = np.pi * md / ng
            win *= (x / np.sin(x)) ** 4
    expect = 2 * (A * ng ** 3 / 2) ** 2 * (1.0 / ng ** 3) * (rl / ng) * win
    assert abs(pk[i] * cnt[i] - expect) < 1e-9 * expect


# ------------------------------------------------------------- growth ----
def test_growth_against_quad(orc):
    g = json.load(open(os.path.join(GOLDEN, "cosmo_golden.json")))
    om = g["cosmology"]["omega_matter"]
    for e in g["growth"]:
        D, dD = orc.growth_f64(e["a"], om, 1 - om)
        assert abs(D / e["D"] - 1) < 1e-9 and abs(dD / e["dDda"] - 1) < 1e-6
    for e in g["growth_eds"]:  # Einstein-de Sitter: D = a exactly
        D, dD = orc.growth_f64(e["a"], 1.0, 0.0)
        assert abs(D - e["a"]) < 1e-10 * e["a"] and abs(dD - 1) < 1e-9


def test_ic_displacement_is_the_potential_flow_of_the_scaled_noise(orc):
    """Zel'dovich IC: for a power-law table (exact under log-log interpolation) the
    displacement must be psi_d^(k) = -i k_d delta^(k) / k^2 with delta^ = FFT(noise) *
    sqrt(P(|k|) N / rl^3), checked with numpy's FFT mode by mode (Nyquist planes
    excluded: the real part taken there breaks Hermitian symmetry), and the velocity
    must be one constant times the displacement."""
    n, rl, seed = 16, 32.0, 7
    A, ns, kd = 3.0, -1.5, 0.01
    table = A * (kd * np.arange(1, 4001)) ** ns
    table = np.concatenate([[0.0], table])
    pos, vel = orc.ic_f64(n, n, rl, seed, table, kd, 0.7, 1.3, 50.0, 2.1, 0.01)
    q = np.indices((n,) * 3).reshape(3, -1).T.astype(np.float64)
    psi = pos - q
    psi -= n * np.round(psi / n)
    noise = orc.white_noise(seed, n ** 3).reshape((n,) * 3)
    m = np.fft.fftfreq(n, 1.0 / n)
    m[m == n // 2] = -n // 2
    k1 = 2 * np.pi / rl * m
    KX, KY, KZ = np.meshgrid(k1, k1, k1, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    P = np.where(k2 > 0, A * np.sqrt(np.where(k2 > 0, k2, 1.0)) ** ns, 0.0) * n ** 3 / rl ** 3
    dhat = np.fft.fftn(noise) * np.sqrt(P)
    inner = (np.abs(m) < n // 2)
    sel = inner[:, None, None] & inner[None, :, None] & inner[None, None, :]
    for d, K in enumerate((KX, KY, KZ)):
        got = np.fft.fftn(psi[:, d].reshape((n,) * 3))
        want = np.where(k2 > 0, -1j * K * dhat / np.where(k2 > 0, k2, 1.0), 0.0)
        err = np.abs(got - want)[sel].max() / np.abs(want[sel]).max()
        assert err < 1e-6, (d, err)
    c = float(np.sum(vel * psi) / np.sum(psi * psi))
    assert np.abs(vel - c * psi).max() < 1e-12 * np.abs(vel).max() + 1e-300


def test_reference_float_growth_close_to_fp64(orc):
    # the reference's 100-bin FP32 trapezoid agrees with the FP64 growth to ~1e-3
    om = 0.31
    for a in (0.02, 0.1, 0.5, 1.0):
        Df = orc.L().orc_ref_delta(a, om, 1 - om)
        D, _ = orc.growth_f64(a, om, 1 - om)
        assert abs(Df / D - 1) < 5e-3


# ---------------------------------------------------------------- runs ----
@pytest.mark.slow
def test_oracle_modes_agree_on_short_run(orc, hg):
    p = hg.make_params(32, 32, 64.0, 50.0, 0.0, 100, seed=11)
    hdr = np.zeros(20)
    ks = np.linspace(0, 50, 5001)
    # EH spectrum from the engine's host writer
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "eh.ipk")
        hg.write_ipk_eh(p, fn, 50.0, 5001)
        hdr, vals = hg.read_ipk(fn)
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    a = 1 / (1 + p.z_ini)
    D, _ = orc.growth_f64(a, p.omega_matter, 1 - p.omega_matter)
    _, dD = orc.growth_f64(a - ts.deltaT / 2, p.omega_matter, 1 - p.omega_matter)
    pos, vel = orc.ic_f64(32, 32, 64.0, 11, vals, hdr[3], D, dD, p.z_ini, ts.fscal, ts.deltaT)
    co = hg.schedule_array(p, 8, merge=False)
    p64, v64, _ = orc.run_f64(32, 32, co, pos, vel)
    cell, off = orc.split_positions(32, pos)
    c2, o2, v2, _ = orc.run_split(32, 32, co, cell, off, vel.astype(np.float32))
    x2 = c2 + o2.astype(np.float64)
    d = (x2 - p64 + 16) % 32 - 16
    assert np.sqrt((d ** 2).mean()) < 1e-4
    assert np.abs(v2 - v64).max() < 1e-3 * np.abs(v64).max()
