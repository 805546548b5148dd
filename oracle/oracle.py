"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The engine never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")
F64, SPLIT, REF32 = 0, 1, 2
_L = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def L():
    global _L
    if _L is None:
        if not os.path.exists(SO):
            build()
        lib = ctypes.CDLL(SO)
        vp, ci, cd, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
        sig = {
            "orc_philox4x32_10": (None, [vp, vp, vp]),
            "orc_det_log": (cd, [cd]),
            "orc_det_cos2pi": (cd, [cd]),
            "orc_normal": (cd, [ctypes.c_uint32, ctypes.c_uint64]),
            "orc_white_noise": (None, [ctypes.c_uint32, i64, vp]),
            "orc_fft3d": (None, [vp, ci, ci]),
            "orc_adot": (cd, [cd] * 9),
            "orc_fscal": (cd, [cd, cd, cd]),
            "orc_ref_delta": (ctypes.c_float, [ctypes.c_float] * 3),
            "orc_ref_dot_delta": (ctypes.c_float, [ctypes.c_float] * 4),
            "orc_growth_f64": (None, [cd, cd, cd, vp, vp]),
            "orc_cic_split": (None, [ci, i64, vp, vp, ctypes.c_float, vp]),
            "orc_cic_f64": (None, [ci, i64, vp, cd, vp]),
            "orc_cic_ref32": (None, [ci, i64, vp, ctypes.c_float, vp]),
            "orc_solve_gradient": (None, [ci, ci, vp, vp]),
            "orc_kick_split": (None, [ci, i64, vp, vp, vp, vp, cd, cd]),
            "orc_kick_f64": (None, [ci, i64, vp, vp, vp, cd, cd]),
            "orc_kick_ref32": (None, [ci, i64, vp, vp, vp, cd, cd]),
            "orc_drift_split": (None, [ci, i64, vp, vp, vp, ctypes.c_float]),
            "orc_drift_f64": (None, [ci, i64, vp, vp, cd]),
            "orc_drift_ref32": (None, [ci, i64, vp, vp, ctypes.c_float]),
            "orc_pk_bins": (None, [ci, cd, ci, ci, vp, vp, vp]),
            "orc_pk_bin_index": (ci, [ci, ci, ci, ci, cd, cd, cd]),
            "orc_power_grid": (None, [ci, cd, ci, ci, ci, vp, vp, vp, vp]),
            "orc_fold_split": (None, [ci, i64, vp, vp]),
            "orc_fold_f64": (None, [ci, i64, vp]),
            "orc_pk_interp": (cd, [ci, ci, ci, vp, cd, cd, ci]),
            "orc_ic_f64": (None, [ci, ci, cd, ctypes.c_uint32, vp, ci, cd, cd, cd, cd, cd, cd, vp, vp]),
            "orc_split_positions": (None, [ci, i64, vp, vp, vp]),
            "orc_run_f64": (cd, [ci, ci, ci, vp, vp, vp, vp]),
            "orc_run_split": (cd, [ci, ci, ci, vp, vp, vp, vp, vp]),
            "orc_run_ref32": (cd, [ci, ci, ci, vp, vp, vp, vp]),
            "orc_num_threads": (ci, []),
            "orc_set_num_threads": (None, [ci]),
        }
        for k, (r, a) in sig.items():
            f = getattr(lib, k)
            f.restype = r
            f.argtypes = a
        _L = lib
    return _L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox(ctr, key):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    o = np.zeros(4, np.uint32)
    L().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o


def white_noise(seed, n):
    out = np.zeros(n, np.float64)
    L().orc_white_noise(seed, n, _p(out))
    return out


def fft3d(grid, sign):
    """Unnormalised 3-D C2C of a complex (ng,ng,ng) array; sign -1 forward, +1 inverse."""
    g = np.ascontiguousarray(grid, np.complex128).copy()
    L().orc_fft3d(_p(g), g.shape[0], int(sign))
    return g


def cic_split(ng, cell, off, mass=1.0):
    cell = np.ascontiguousarray(cell, np.int32)
    off = np.ascontiguousarray(off, np.float32)
    out = np.zeros(ng ** 3, np.int64)
    L().orc_cic_split(ng, cell.size // 3, _p(cell), _p(off), float(mass), _p(out))
    return out.reshape(ng, ng, ng)


def cic_f64(ng, pos, mass=1.0):
    pos = np.ascontiguousarray(pos, np.float64)
    out = np.zeros(ng ** 3, np.float64)
    L().orc_cic_f64(ng, pos.size // 3, _p(pos), float(mass), _p(out))
    return out.reshape(ng, ng, ng)


def cic_ref32(ng, pos, mass=1.0):
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.zeros(ng ** 3, np.float32)
    L().orc_cic_ref32(ng, pos.size // 3, _p(pos), float(mass), _p(out))
    return out.reshape(ng, ng, ng)


def solve_gradient(ng, rho, mode=F64):
    rho = np.ascontiguousarray(rho, np.float64).reshape(-1)
    f = np.zeros((ng ** 3, 3), np.float64)
    L().orc_solve_gradient(ng, mode, _p(rho), _p(f))
    return f


def kick_split(ng, cell, off, vel, force, deltaT, fscal):
    cell = np.ascontiguousarray(cell, np.int32)
    off = np.ascontiguousarray(off, np.float32)
    v = np.ascontiguousarray(vel, np.float32).copy()
    force = np.ascontiguousarray(force, np.float32)
    L().orc_kick_split(ng, cell.size // 3, _p(cell), _p(off), _p(v), _p(force), float(deltaT), float(fscal))
    return v


def kick_f64(ng, pos, vel, force, deltaT, fscal):
    pos = np.ascontiguousarray(pos, np.float64)
    v = np.ascontiguousarray(vel, np.float64).copy()
    force = np.ascontiguousarray(force, np.float64)
    L().orc_kick_f64(ng, pos.size // 3, _p(pos), _p(v), _p(force), float(deltaT), float(fscal))
    return v


def kick_ref32(ng, pos, vel, force, deltaT, fscal):
    pos = np.ascontiguousarray(pos, np.float32)
    v = np.ascontiguousarray(vel, np.float32).copy()
    force = np.ascontiguousarray(force, np.float32)
    L().orc_kick_ref32(ng, pos.size // 3, _p(pos), _p(v), _p(force), float(deltaT), float(fscal))
    return v


def drift_ref32(ng, pos, vel, pf):
    p = np.ascontiguousarray(pos, np.float32).copy()
    v = np.ascontiguousarray(vel, np.float32)
    L().orc_drift_ref32(ng, p.size // 3, _p(p), _p(v), float(pf))
    return p


def drift_split(ng, cell, off, vel, pf):
    c = np.ascontiguousarray(cell, np.int32).copy()
    o = np.ascontiguousarray(off, np.float32).copy()
    v = np.ascontiguousarray(vel, np.float32)
    L().orc_drift_split(ng, c.size // 3, _p(c), _p(o), _p(v), float(pf))
    return c, o


def drift_f64(ng, pos, vel, pf):
    p = np.ascontiguousarray(pos, np.float64).copy()
    v = np.ascontiguousarray(vel, np.float64)
    L().orc_drift_f64(ng, p.size // 3, _p(p), _p(v), float(pf))
    return p


def pk_bins(ng, rl, folds, nbins):
    mk, bd = ctypes.c_double(), ctypes.c_double()
    kb = np.zeros(nbins)
    L().orc_pk_bins(ng, float(rl), folds, nbins, ctypes.byref(mk), ctypes.byref(bd), _p(kb))
    return mk.value, bd.value, kb


def power_grid(ng, rl, folds, fold, nbins, rho):
    rho = np.ascontiguousarray(rho, np.float64).reshape(-1)
    k = np.zeros(nbins)
    pk = np.zeros(nbins)
    cnt = np.zeros(nbins, np.int64)
    L().orc_power_grid(ng, float(rl), folds, fold, nbins, _p(rho), _p(k), _p(pk), _p(cnt))
    return k, pk, cnt


def fold_split(ng, cell, off):
    c = np.ascontiguousarray(cell, np.int32).copy()
    o = np.ascontiguousarray(off, np.float32).copy()
    L().orc_fold_split(ng, c.size // 3, _p(c), _p(o))
    return c, o


def fold_f64(ng, pos):
    p = np.ascontiguousarray(pos, np.float64).copy()
    L().orc_fold_f64(ng, p.size // 3, _p(p))
    return p


def ic_f64(ng, np_, rl, seed, ipk_vals, k_delta, delta, dot_delta, z_ini, fscal, deltaT):
    v = np.ascontiguousarray(ipk_vals, np.float64)
    n = np_ ** 3
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    L().orc_ic_f64(ng, np_, float(rl), seed, _p(v), v.size, float(k_delta), float(delta), float(dot_delta),
                   float(z_ini), float(fscal), float(deltaT), _p(pos), _p(vel))
    return pos, vel


def split_positions(ng, pos):
    pos = np.ascontiguousarray(pos, np.float64)
    cell = np.zeros(pos.shape, np.int32)
    off = np.zeros(pos.shape, np.float32)
    L().orc_split_positions(ng, pos.size // 3, _p(pos), _p(cell), _p(off))
    return cell, off


STAGES = ("drift", "deposit", "solve", "gather_kick")


def run_f64(ng, np_, coeffs, pos, vel, stages=None):
    """stages: a dict that receives the seconds of each step stage (STAGES)."""
    c = np.ascontiguousarray(coeffs, np.float64)
    p = np.ascontiguousarray(pos, np.float64).copy()
    v = np.ascontiguousarray(vel, np.float64).copy()
    st = np.zeros(4)
    t = L().orc_run_f64(ng, np_, c.shape[0], _p(c), _p(p), _p(v), _p(st))
    if stages is not None:
        stages.update(zip(STAGES, st.tolist()))
    return p, v, t


def run_split(ng, np_, coeffs, cell, off, vel, stages=None):
    c = np.ascontiguousarray(coeffs, np.float64)
    ce = np.ascontiguousarray(cell, np.int32).copy()
    o = np.ascontiguousarray(off, np.float32).copy()
    v = np.ascontiguousarray(vel, np.float32).copy()
    st = np.zeros(4)
    t = L().orc_run_split(ng, np_, c.shape[0], _p(c), _p(ce), _p(o), _p(v), _p(st))
    if stages is not None:
        stages.update(zip(STAGES, st.tolist()))
    return ce, o, v, t


def run_ref32(ng, np_, coeffs, pos, vel, stages=None):
    c = np.ascontiguousarray(coeffs, np.float64)
    p = np.ascontiguousarray(pos, np.float32).copy()
    v = np.ascontiguousarray(vel, np.float32).copy()
    st = np.zeros(4)
    t = L().orc_run_ref32(ng, np_, c.shape[0], _p(c), _p(p), _p(v), _p(st))
    if stages is not None:
        stages.update(zip(STAGES, st.tolist()))
    return p, v, t


def growth_f64(a, om, ol):
    d, dd = ctypes.c_double(), ctypes.c_double()
    L().orc_growth_f64(float(a), float(om), float(ol), ctypes.byref(d), ctypes.byref(dd))
    return d.value, dd.value


def set_threads(n):
    L().orc_set_num_threads(int(n))


def num_threads():
    return L().orc_num_threads()
