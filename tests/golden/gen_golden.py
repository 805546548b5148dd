"""Writes the golden fixtures under tests/golden/ (run once, outputs committed).

- fft_golden.npz: seeded random complex grids (8^3, 12^3, 16^3) and numpy.fft.fftn
  of each (forward, e^{-i}, unnormalised) -- pins the 3-D C2C transform.
- cosmo_golden.json: an independent numpy/scipy statement of
  * the Eisenstein & Hu 1998 no-wiggle transfer function (ApJ 496, 605, eqs. 26, 28-31),
  * sigma_8 normalisation of A k^ns T^2 (scipy.integrate.quad, top hat R = 8 Mpc/h),
  * the linear growth D(a) = 5/2 Om E(a) int_0^a da'/(a'E)^3 and dD/da (quad),
  for the cosmology of the measured configurations.
- test_run_params / test_run_ipk.ipk: the reference's example params file and .ipk
  table, copied as data (the params parser and .ipk reader read them).

The reference's own Python is not imported: the fixtures come from numpy/scipy.
"""
import json
import os
import shutil

import numpy as np
from scipy import integrate

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_fixtures():
    out = {}
    for n in (8, 12, 16):
        rng = np.random.default_rng(1000 + n)
        x = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
        out[f"x{n}"] = x
        out[f"X{n}"] = np.fft.fftn(x)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


def eh_nowiggle(k, om, ob, h, tcmb):
    theta = tcmb / 2.7
    omh2, obh2, fb = om * h * h, ob * h * h, ob / om
    s = 44.5 * np.log(9.83 / omh2) / np.sqrt(1 + 10 * obh2 ** 0.75)
    ag = 1 - 0.328 * np.log(431 * omh2) * fb + 0.38 * np.log(22.3 * omh2) * fb ** 2
    gam = om * h * (ag + (1 - ag) / (1 + (0.43 * k * h * s) ** 4))
    q = k * theta ** 2 / gam
    L0 = np.log(2 * np.e + 1.8 * q)
    C0 = 14.2 + 731 / (1 + 62.5 * q)
    return L0 / (L0 + C0 * q * q)


def sigma2(R, ns, om, ob, h, tcmb):
    def f(lnk):
        k = np.exp(lnk)
        x = k * R
        W = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3 if x > 1e-3 else 1 - x * x / 10
        return k ** 3 * k ** ns * eh_nowiggle(k, om, ob, h, tcmb) ** 2 * W * W / (2 * np.pi ** 2)
    val, _ = integrate.quad(f, np.log(1e-6), np.log(1e4), limit=2000, epsabs=0, epsrel=1e-11)
    return val


def growth(a, om, ol):
    aE2 = lambda x: 1 + om * (1 / x - 1) + ol * (x * x - 1)
    I, _ = integrate.quad(lambda x: aE2(x) ** -1.5 if x > 0 else 0.0, 0, a, epsabs=0, epsrel=1e-13, limit=500)
    q = aE2(a)
    E = np.sqrt(q) / a
    D = 2.5 * om * E * I
    eps = 1e-5 * a
    def Dof(b):
        Ib, _ = integrate.quad(lambda x: aE2(x) ** -1.5 if x > 0 else 0.0, 0, b, epsabs=0, epsrel=1e-13, limit=500)
        return 2.5 * om * np.sqrt(aE2(b)) / b * Ib
    dD = (Dof(a + eps) - Dof(a - eps)) / (2 * eps)
    return D, dD


def cosmo_fixtures():
    f32 = lambda x: float(np.float32(x))
    h = f32(0.68)
    ocdm, deut = f32(0.261), f32(0.0226576)
    ob = deut / h / h
    om = ocdm + ob
    tcmb, ns, s8 = f32(2.726), f32(0.96), f32(0.81)
    ks = [1e-4, 1e-3, 0.01, 0.05, 0.1, 0.3, 1.0, 3.0, 10.0, 100.0]
    T = [float(eh_nowiggle(k, om, ob, h, tcmb)) for k in ks]
    s2 = sigma2(8.0, ns, om, ob, h, tcmb)
    A = s8 * s8 / s2
    ol = 1 - om
    grow = []
    for a in (1.0 / 201, 1.0 / 51, 1.0 / 31, 1.0 / 41, 0.5, 1.0):
        D, dD = growth(a, om, ol)
        grow.append({"a": a, "D": D, "dDda": dD})
    eds = []
    for a in (0.01, 0.1, 0.5, 1.0):
        D, dD = growth(a, 1.0, 0.0)
        eds.append({"a": a, "D": D, "dDda": dD})
    D0, _ = growth(1.0, om, ol)
    zi = f32(50.0)
    Di, _ = growth(1.0 / (1.0 + zi), om, ol)
    pk_z = [A * k ** ns * float(eh_nowiggle(k, om, ob, h, tcmb)) ** 2 * (Di / D0) ** 2 for k in (0.05, 0.1, 0.5, 1.0)]
    out = {
        "cosmology": {"omega_cdm": ocdm, "deut": deut, "hubble": h, "omega_baryon": ob, "omega_matter": om,
                      "t_cmb": tcmb, "ns": ns, "ss8": s8},
        "eh_k": ks, "eh_T": T, "sigma2_unit": s2, "A": A,
        "growth": grow, "growth_eds": eds,
        "pk_z50": {"z": zi, "k": [0.05, 0.1, 0.5, 1.0], "P": pk_z},
    }
    with open(os.path.join(HERE, "cosmo_golden.json"), "w") as f:
        json.dump(out, f, indent=1)


def reference_data(ref="/root/reference"):
    for name in ("test_run_params", "test_run_ipk.ipk"):
        src = os.path.join(ref, name)
        if os.path.exists(src):
            shutil.copyfile(src, os.path.join(HERE, name))


if __name__ == "__main__":
    fft_fixtures()
    cosmo_fixtures()
    reference_data()
    print("fixtures written to", HERE)
