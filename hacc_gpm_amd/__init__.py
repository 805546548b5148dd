"""hacc_gpm_amd — MI355X particle-mesh engine (drop-in for HACC_GPM's serial PM loop).

Python mirror of the reference's operator interface (namespace HACCGPM::serial,
src/haccgpm.hpp:385-506) over the C ABI of libhgpm.so (include/hgpm.h):

    params = read_params("run.params")          # HACCGPM::read_params
    ts = Timestepper(params)                    # HACCGPM::Timestepper
    mem = MemoryManager(params)                 # serial::MemoryManager (+ fft plans)
    GenerateDisplacementIC(mem, params, ts, ipk)
    UpdatePositions(params, mem, ts, 0.5); CIC(params, mem); SolveGradient(params, mem)
    UpdateVelocities(params, mem, ts); GetPowerSpectrum(params, mem, "out.pk.0")

Every call goes to the HIP engine; there is no CPU fallback.  Importing works
without a GPU (host-side functions such as read_params, the timestepper and the
.ipk reader run on the CPU); creating a MemoryManager needs a GPU and raises
HgpmError otherwise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhgpm.so")
DRIVER_PATH = os.path.join(_HERE, "bin", "haccgpm")

OK, ERR_ARG, ERR_HIP, ERR_ALLOC, ERR_IO, ERR_UNSUPPORTED, ERR_STATE, ERR_NUMERIC = range(8)
FP32, FP64 = 0, 1
MAX_STEPS = 1000
IC_STATIC_BINS = 262144  # HGPM_IC_STATIC_BINS: largest .ipk table of generate_ic_noalloc
MASS_UNIT = 2.0 ** -24


class HgpmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hgpm status {code}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """HACCGPM::Params (src/haccgpm.hpp:118-174) as laid out in include/hgpm.h."""
    _fields_ = [
        ("fname", ctypes.c_char * 300),
        ("ng", ctypes.c_int), ("np", ctypes.c_int),
        ("rl", ctypes.c_double),
        ("block_size", ctypes.c_int), ("seed", ctypes.c_int),
        ("z_ini", ctypes.c_double), ("z_fin", ctypes.c_double),
        ("nsteps", ctypes.c_int), ("last_step", ctypes.c_int),
        ("prefix", ctypes.c_char * 300), ("ipk", ctypes.c_char * 300),
        ("hubble", ctypes.c_double), ("omega_cdm", ctypes.c_double), ("omega_nu", ctypes.c_double),
        ("deut", ctypes.c_double), ("ss8", ctypes.c_double), ("ns", ctypes.c_double),
        ("w_de", ctypes.c_double), ("wa_de", ctypes.c_double), ("t_cmb", ctypes.c_double),
        ("neff_massless", ctypes.c_double), ("neff_massive", ctypes.c_double),
        ("omega_baryon", ctypes.c_double), ("omega_cb", ctypes.c_double),
        ("omega_matter", ctypes.c_double), ("omega_radiation", ctypes.c_double),
        ("f_nu_massless", ctypes.c_double), ("f_nu_massive", ctypes.c_double),
        ("pk_folds", ctypes.c_int), ("pk_bins", ctypes.c_int),
        ("pk_dump", ctypes.c_ubyte * MAX_STEPS), ("particle_dump", ctypes.c_ubyte * MAX_STEPS),
        ("dump_init", ctypes.c_int), ("dump_final", ctypes.c_int),
        ("ol", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("analysis", ctypes.c_ubyte * MAX_STEPS), ("analysis_dir", ctypes.c_char * 300),
        ("analysis_py", ctypes.c_char * 300), ("do_analysis", ctypes.c_int),
    ]


class TimestepperT(ctypes.Structure):
    _fields_ = [("aa", ctypes.c_double), ("z", ctypes.c_double), ("deltaT", ctypes.c_double),
                ("adot", ctypes.c_double), ("fscal", ctypes.c_double)]


class StepCoeffs(ctypes.Structure):
    _fields_ = [("drift0", ctypes.c_double), ("kick_dt", ctypes.c_double), ("kick_fscal", ctypes.c_double),
                ("drift1", ctypes.c_double), ("skip_drift0", ctypes.c_int)]


_lib = None


def lib():
    """Loads libhgpm.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HgpmError(ERR_IO, f"{LIB_PATH} is missing: run `python -m hacc_gpm_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        c_int, c_double, c_char_p, vp = ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_void_p
        sig = {
            "hgpm_version": (c_char_p, []),
            "hgpm_last_error": (c_char_p, [vp]),
            "hgpm_device_count": (c_int, [P(c_int)]),
            "hgpm_read_params": (c_int, [c_char_p, P(Params)]),
            "hgpm_save_params": (c_int, [P(Params)]),
            "hgpm_read_ipk": (c_int, [c_char_p, vp, vp, c_int, P(c_int)]),
            "hgpm_write_ipk_eh": (c_int, [P(Params), c_char_p, c_double, c_int]),
            "hgpm_get_adot": (c_double, [P(Params), c_double]),
            "hgpm_get_fscal": (c_double, [c_double, c_double, c_double]),
            "hgpm_ts_init": (c_int, [P(Params), P(TimestepperT)]),
            "hgpm_ts_set_initial_z": (c_int, [P(Params), P(TimestepperT), c_double]),
            "hgpm_ts_advance_half_step": (c_int, [P(Params), P(TimestepperT)]),
            "hgpm_ts_reverse_half_step": (c_int, [P(Params), P(TimestepperT)]),
            "hgpm_schedule": (c_int, [P(Params), c_int, c_int, vp, P(StepCoeffs)]),
            "hgpm_growth": (c_int, [P(Params), c_double, P(c_double), P(c_double)]),
            "hgpm_create": (c_int, [P(Params), P(vp)]),
            "hgpm_destroy": (c_int, [vp]),
            "hgpm_device_bytes": (c_int, [vp, P(ctypes.c_size_t)]),
            "hgpm_synchronize": (c_int, [vp]),
            "hgpm_set_particles": (c_int, [vp, vp, vp]),
            "hgpm_get_particles": (c_int, [vp, vp, vp]),
            "hgpm_set_particles_split": (c_int, [vp, vp, vp, vp]),
            "hgpm_get_particles_split": (c_int, [vp, vp, vp, vp]),
            "hgpm_set_particles_f64": (c_int, [vp, vp, vp]),
            "hgpm_get_particles_f64": (c_int, [vp, vp, vp]),
            "hgpm_copy_particles": (c_int, [vp, vp]),
            "hgpm_analysis_particles": (c_int, [vp, vp]),
            "hgpm_compare_particles": (c_int, [vp, vp, vp]),
            "hgpm_generate_ic": (c_int, [vp, vp, c_int, c_double, c_double, c_double, c_double, c_double, c_double]),
            "hgpm_autotune_particles": (c_int, [vp, c_int, vp, vp]),
            "hgpm_generate_ic_noalloc": (c_int, [vp, vp, c_int, c_double, c_double, c_double, c_double, c_double,
                                                 c_double]),
            "hgpm_update_positions": (c_int, [vp, c_double]),
            "hgpm_cic": (c_int, [vp]),
            "hgpm_get_density": (c_int, [vp, vp]),
            "hgpm_get_density_f64": (c_int, [vp, vp]),
            "hgpm_set_density": (c_int, [vp, vp]),
            "hgpm_set_density_f64": (c_int, [vp, vp]),
            "hgpm_solve_gradient": (c_int, [vp]),
            "hgpm_get_force": (c_int, [vp, vp]),
            "hgpm_get_force_f64": (c_int, [vp, vp]),
            "hgpm_update_velocities": (c_int, [vp, c_double, c_double]),
            "hgpm_kick_drift": (c_int, [vp, c_double, c_double, c_double]),
            "hgpm_step": (c_int, [vp, P(StepCoeffs)]),
            "hgpm_fft_host": (c_int, [vp, vp, c_int]),
            "hgpm_fft_device": (c_int, [vp, vp, c_int]),
            "hgpm_power_spectrum": (c_int, [vp, c_int, vp, vp, vp]),
            "hgpm_power_spectrum_grid": (c_int, [vp, c_int, vp, vp, vp]),
            "hgpm_write_power_spectrum": (c_int, [vp, c_char_p]),
            "hgpm_write_output": (c_int, [vp, c_char_p]),
            "hgpm_set_timing": (c_int, [vp, c_int]),
            "hgpm_get_timings": (c_int, [vp, vp, c_int, vp, vp, c_int, P(c_int)]),
            "hgpm_reset_timings": (c_int, [vp]),
            "hgpm_set_tunable": (c_int, [vp, c_char_p, c_int]),
            "hgpm_get_tunable": (c_int, [vp, c_char_p, P(c_int)]),
            "hgpm_autotune": (c_int, [vp, c_int, vp, vp]),
            "hgpm_ensemble_create": (c_int, [c_int, vp, P(vp)]),
            "hgpm_ensemble_destroy": (c_int, [vp]),
            "hgpm_ensemble_last_error": (c_char_p, [vp]),
            "hgpm_ensemble_create_budget": (c_int, [c_int, vp, ctypes.c_size_t, P(vp)]),
            "hgpm_ensemble_waves": (c_int, [vp, vp]),
            "hgpm_ensemble_batches": (c_int, [vp, P(c_int), vp]),
            "hgpm_ensemble_device_bytes": (c_int, [vp, P(ctypes.c_size_t)]),
            "hgpm_ensemble_generate_ic": (c_int, [vp, c_int, vp, c_int, c_double, c_double, c_double, c_double,
                                                  c_double, c_double]),
            "hgpm_ensemble_set_particles": (c_int, [vp, c_int, vp, vp]),
            "hgpm_ensemble_get_particles": (c_int, [vp, c_int, vp, vp]),
            "hgpm_ensemble_power_spectrum": (c_int, [vp, c_int, c_int, vp, vp, vp]),
            "hgpm_ensemble_step": (c_int, [vp, vp]),
            "hgpm_ensemble_run": (c_int, [vp, c_int, vp]),
            "hgpm_ensemble_synchronize": (c_int, [vp]),
            "hgpm_ensemble_autotune": (c_int, [vp, c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names declared in include/hgpm.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "hgpm.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(hgpm_[a-z0-9_]+)\s*\(", txt)))


def _check(rc, ctx=None):
    if rc != OK:
        msg = lib().hgpm_last_error(ctx)
        raise HgpmError(rc, msg.decode() if msg else "")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def version():
    return lib().hgpm_version().decode()


def device_count():
    """HIP devices visible to this process (0 without a GPU), asked of libhgpm's own HIP
    runtime, so no other runtime is initialised in this process."""
    n = ctypes.c_int(0)
    _check(lib().hgpm_device_count(ctypes.byref(n)))
    return n.value


# ----------------------------------------------------------- host side ----
def read_params(fname):
    """HACCGPM::read_params (src/simulation/params.cpp:72-251)."""
    p = Params()
    _check(lib().hgpm_read_params(str(fname).encode(), ctypes.byref(p)))
    return p


def save_params(p):
    _check(lib().hgpm_save_params(ctypes.byref(p)))


def read_ipk(fname):
    """CosmoClass::read_ipk (cosmo.cpp:26-79): returns (header[20], values)."""
    nb = ctypes.c_int(0)
    hdr = np.zeros(20, np.float64)
    _check(lib().hgpm_read_ipk(str(fname).encode(), _ptr(hdr), None, 0, ctypes.byref(nb)))
    vals = np.zeros(nb.value, np.float64)
    _check(lib().hgpm_read_ipk(str(fname).encode(), _ptr(hdr), _ptr(vals), nb.value, ctypes.byref(nb)))
    return hdr, vals


def write_ipk_eh(p, fname, k_max=1000.0, k_bins=20000):
    _check(lib().hgpm_write_ipk_eh(ctypes.byref(p), str(fname).encode(), float(k_max), int(k_bins)))


def get_adot(p, a):
    return lib().hgpm_get_adot(ctypes.byref(p), float(a))


def get_fscal(a, adot, omega_cb):
    return lib().hgpm_get_fscal(float(a), float(adot), float(omega_cb))


def growth(p, a):
    d, dd = ctypes.c_double(), ctypes.c_double()
    _check(lib().hgpm_growth(ctypes.byref(p), float(a), ctypes.byref(d), ctypes.byref(dd)))
    return d.value, dd.value


def schedule(p, n, merge=True, output_after=None):
    co = (StepCoeffs * max(n, 1))()
    oa = None
    if output_after is not None:
        oa_arr = np.asarray(output_after, np.uint8)
        oa = _ptr(oa_arr)
    _check(lib().hgpm_schedule(ctypes.byref(p), int(n), 1 if merge else 0, oa, co))
    return [co[i] for i in range(n)]


def schedule_array(p, n, merge=True, output_after=None):
    """(n, 4) float64: drift0 (0 when merged away), kick_dt, kick_fscal, drift1."""
    out = np.zeros((n, 4))
    for i, c in enumerate(schedule(p, n, merge, output_after)):
        out[i] = (0.0 if c.skip_drift0 else c.drift0, c.kick_dt, c.kick_fscal, c.drift1)
    return out


class Timestepper:
    """HACCGPM::Timestepper (src/simulation/timestepper.cpp:31-68)."""

    def __init__(self, params):
        self.params = params
        self.t = TimestepperT()
        _check(lib().hgpm_ts_init(ctypes.byref(params), ctypes.byref(self.t)))

    def setInitialZ(self, z):
        lib().hgpm_ts_set_initial_z(ctypes.byref(self.params), ctypes.byref(self.t), float(z))

    def advanceHalfStep(self):
        lib().hgpm_ts_advance_half_step(ctypes.byref(self.params), ctypes.byref(self.t))

    def reverseHalfStep(self):
        lib().hgpm_ts_reverse_half_step(ctypes.byref(self.params), ctypes.byref(self.t))

    aa = property(lambda s: s.t.aa)
    z = property(lambda s: s.t.z)
    deltaT = property(lambda s: s.t.deltaT)
    adot = property(lambda s: s.t.adot)
    fscal = property(lambda s: s.t.fscal)

    def drift_prefactor(self, frac=0.5):
        """float(dt/(a^2 adot) * frac), UpdatePositions (pm_actions.cpp:291)."""
        return float(np.float32((self.deltaT / (self.aa * self.aa * self.adot)) * np.float32(frac)))


# ------------------------------------------------------------- context ----
class MemoryManager:
    """serial::MemoryManager (src/simulation/cuda/mmanager.cu:93-229): owns the device
    buffers of one box.  All methods run on the GPU through libhgpm.so."""

    def __init__(self, params):
        self.params = params
        self.ng, self.np = params.ng, params.np
        self.N = params.np ** 3
        self.G = params.ng ** 3
        self.precision = params.precision
        h = ctypes.c_void_p()
        _check(lib().hgpm_create(ctypes.byref(params), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().hgpm_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _check(rc, self.h)

    def device_bytes(self):
        b = ctypes.c_size_t()
        self._c(lib().hgpm_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def synchronize(self):
        self._c(lib().hgpm_synchronize(self.h))

    # particles
    def set_particles(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1)
        assert pos.size == 3 * self.N and vel.size == 3 * self.N
        self._c(lib().hgpm_set_particles(self.h, _ptr(pos), _ptr(vel)))

    def get_particles(self):
        pos = np.zeros((self.N, 3), np.float32)
        vel = np.zeros((self.N, 3), np.float32)
        self._c(lib().hgpm_get_particles(self.h, _ptr(pos), _ptr(vel)))
        return pos, vel

    def set_particles_split(self, cell, off, vel):
        cell = np.ascontiguousarray(cell, np.int32).reshape(-1)
        off = np.ascontiguousarray(off, np.float32).reshape(-1)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1)
        assert cell.size == off.size == vel.size == 3 * self.N
        self._c(lib().hgpm_set_particles_split(self.h, _ptr(cell), _ptr(off), _ptr(vel)))

    def get_particles_split(self):
        cell = np.zeros((self.N, 3), np.int32)
        off = np.zeros((self.N, 3), np.float32)
        vel = np.zeros((self.N, 3), np.float32)
        self._c(lib().hgpm_get_particles_split(self.h, _ptr(cell), _ptr(off), _ptr(vel)))
        return cell, off, vel

    def set_particles_f64(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1)
        vel = np.ascontiguousarray(vel, np.float64).reshape(-1)
        assert pos.size == 3 * self.N and vel.size == 3 * self.N
        self._c(lib().hgpm_set_particles_f64(self.h, _ptr(pos), _ptr(vel)))

    def get_particles_f64(self):
        pos = np.zeros((self.N, 3), np.float64)
        vel = np.zeros((self.N, 3), np.float64)
        self._c(lib().hgpm_get_particles_f64(self.h, _ptr(pos), _ptr(vel)))
        return pos, vel

    def analysis_particles(self):
        """The array serial::PyAnalysis hands to the analysis module
        (src/analysis/cuda/py_analysis.cu:3-26): (np^3, 7) float32 rows of
        (index, x, y, z, vx, vy, vz) in the reference's particle order."""
        out = np.zeros((self.N, 7), np.float32)
        self._c(lib().hgpm_analysis_particles(self.h, _ptr(out)))
        return out

    def copy_particles_from(self, other):
        """Particle state of another context (same NP, NG), converted on the device."""
        self._c(lib().hgpm_copy_particles(self.h, other.h))

    def compare_particles(self, other):
        """RMS periodic position difference (grid units), velocity RMS difference
        relative to other's RMS velocity, and the largest position difference of a
        component, against another context of the same lattice (FP64 sums on the device)."""
        out = np.zeros(4)
        self._c(lib().hgpm_compare_particles(self.h, other.h, _ptr(out)))
        return {"rms_disp_cells": float(np.sqrt(out[0] / self.N)), "vel_rel_rms": float(np.sqrt(out[1] / out[2]))
                if out[2] > 0 else 0.0, "max_disp_cells": float(out[3]), "sums": out.tolist()}

    # steps
    def generate_ic(self, ipk_header, ipk_values, delta, dot_delta, fscal, deltaT):
        v = np.ascontiguousarray(ipk_values, np.float64)
        self._c(lib().hgpm_generate_ic(self.h, _ptr(v), int(v.size), float(ipk_header[3]), float(ipk_header[1]),
                                       float(delta), float(dot_delta), float(fscal), float(deltaT)))

    def generate_ic_noalloc(self, ipk_header, ipk_values, delta, dot_delta, fscal, deltaT):
        """generate_ic with the .ipk table staged in the device's fixed array (no device
        allocation; tables of at most IC_STATIC_BINS bins)."""
        v = np.ascontiguousarray(ipk_values, np.float64)
        self._c(lib().hgpm_generate_ic_noalloc(self.h, _ptr(v), int(v.size), float(ipk_header[3]),
                                               float(ipk_header[1]), float(delta), float(dot_delta), float(fscal),
                                               float(deltaT)))

    def update_positions(self, prefactor):
        self._c(lib().hgpm_update_positions(self.h, float(prefactor)))

    def cic(self):
        self._c(lib().hgpm_cic(self.h))

    def get_density(self):
        r = np.zeros(self.G, np.int64)
        self._c(lib().hgpm_get_density(self.h, _ptr(r)))
        return r.reshape(self.ng, self.ng, self.ng)

    def get_density_f64(self):
        r = np.zeros(self.G, np.float64)
        self._c(lib().hgpm_get_density_f64(self.h, _ptr(r)))
        return r.reshape(self.ng, self.ng, self.ng)

    def set_density(self, rho):
        r = np.ascontiguousarray(rho, np.int64).reshape(-1)
        assert r.size == self.G
        self._c(lib().hgpm_set_density(self.h, _ptr(r)))

    def set_density_f64(self, rho):
        r = np.ascontiguousarray(rho, np.float64).reshape(-1)
        assert r.size == self.G
        self._c(lib().hgpm_set_density_f64(self.h, _ptr(r)))

    def solve_gradient(self):
        self._c(lib().hgpm_solve_gradient(self.h))

    def get_force(self):
        if self.precision == FP64:
            f = np.zeros((self.G, 3), np.float64)
            self._c(lib().hgpm_get_force_f64(self.h, _ptr(f)))
        else:
            f = np.zeros((self.G, 3), np.float32)
            self._c(lib().hgpm_get_force(self.h, _ptr(f)))
        return f

    def update_velocities(self, deltaT, fscal):
        self._c(lib().hgpm_update_velocities(self.h, float(deltaT), float(fscal)))

    def kick_drift(self, deltaT, fscal, prefactor):
        self._c(lib().hgpm_kick_drift(self.h, float(deltaT), float(fscal), float(prefactor)))

    def step(self, coeffs):
        self._c(lib().hgpm_step(self.h, ctypes.byref(coeffs)))

    def fft(self, grid, direction):
        """In-place unnormalised 3-D C2C (forward_fft / backward_fft, ffts.cu)."""
        dt = np.complex64 if self.precision == FP32 else np.complex128
        g = np.ascontiguousarray(grid, dt).copy()
        assert g.size == self.G
        self._c(lib().hgpm_fft_host(self.h, _ptr(g), int(direction)))
        return g.reshape(grid.shape)

    def power_spectrum(self, fold=0, from_grid=False):
        nb = self.params.pk_bins
        k = np.zeros(nb)
        pk = np.zeros(nb)
        cnt = np.zeros(nb, np.int64)
        f = lib().hgpm_power_spectrum_grid if from_grid else lib().hgpm_power_spectrum
        self._c(f(self.h, int(fold), _ptr(k), _ptr(pk), _ptr(cnt)))
        return k, pk, cnt

    def write_power_spectrum(self, fname):
        self._c(lib().hgpm_write_power_spectrum(self.h, str(fname).encode()))

    def write_output(self, fname):
        self._c(lib().hgpm_write_output(self.h, str(fname).encode()))

    # tuning
    def set_tunable(self, key, value):
        self._c(lib().hgpm_set_tunable(self.h, key.encode(), int(value)))

    def get_tunable(self, key):
        v = ctypes.c_int()
        self._c(lib().hgpm_get_tunable(self.h, key.encode(), ctypes.byref(v)))
        return v.value

    def autotune(self, reps=3):
        """Times the FFT variants on this grid and keeps the fastest, then, when the context
        has particles, the deposit tile and gather shapes; returns the per-variant
        milliseconds (column pass, fused x pass of the solve, deposit, gather)."""
        n = self.get_tunable("fft_variants")
        mc = np.zeros(n)
        mx = np.zeros(n)
        self._c(lib().hgpm_autotune(self.h, int(reps), _ptr(mc), _ptr(mx)))
        out = {"cols_ms": [v if v >= 0 else None for v in mc.tolist()],
               "xsolve_ms": [v if v >= 0 else None for v in mx.tolist()],
               "fft_cols": self.get_tunable("fft_cols"), "fft_xsolve": self.get_tunable("fft_xsolve"),
               "fft_plan": self.get_tunable("fft_plan")}
        if self.precision == FP32:
            # the particle kernels, on this context's particles when it has them
            md, mg = np.zeros(2), np.zeros(2)
            rc = lib().hgpm_autotune_particles(self.h, int(reps), _ptr(md), _ptr(mg))
            if rc == OK:
                out.update({"deposit_ms": md.tolist(), "gather_ms": mg.tolist(),
                            "deposit_tile": self.get_tunable("deposit_tile"),
                            "gather_tile": self.get_tunable("gather_tile")})
            elif rc != ERR_STATE:
                self._c(rc)
        return out

    # instrumentation
    def set_timing(self, on=True):
        self._c(lib().hgpm_set_timing(self.h, 1 if on else 0))

    def reset_timings(self):
        self._c(lib().hgpm_reset_timings(self.h))

    def timings(self):
        names = ctypes.create_string_buffer(8192)
        ms = np.zeros(128)
        cnt = np.zeros(128, np.int64)
        n = ctypes.c_int()
        self._c(lib().hgpm_get_timings(self.h, names, 8192, _ptr(ms), _ptr(cnt), 128, ctypes.byref(n)))
        parts = names.raw.split(b"\0")
        return {parts[i].decode(): (float(ms[i]), int(cnt[i])) for i in range(n.value)}


# -------------------------------------------------------------- ensemble ----
class Ensemble:
    """Many boxes stepped together (hgpm_ensemble_*): boxes with the same dynamics are
    batched into one set of launches per step; seeds and spectra stay per box."""

    def __init__(self, params_list, device_budget=0):
        """device_budget: bytes the ensemble may use (0: 90% of the free device memory);
        batches that do not fit keep grid buffers for a wave of their boxes only."""
        self.params = list(params_list)
        arr = (Params * len(self.params))(*self.params)
        h = ctypes.c_void_p()
        rc = lib().hgpm_ensemble_create_budget(len(self.params), arr, int(device_budget), ctypes.byref(h))
        if rc:
            raise HgpmError(rc, lib().hgpm_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().hgpm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        if rc:
            raise HgpmError(rc, lib().hgpm_ensemble_last_error(self.h).decode())

    def batches(self):
        n = ctypes.c_int()
        of = np.zeros(len(self.params), np.int32)
        self._c(lib().hgpm_ensemble_batches(self.h, ctypes.byref(n), _ptr(of)))
        return n.value, of

    def waves(self):
        """Boxes per wave of each batch (a batch's size when it is fully resident)."""
        n, _ = self.batches()
        w = np.zeros(n, np.int32)
        self._c(lib().hgpm_ensemble_waves(self.h, _ptr(w)))
        return w

    def device_bytes(self):
        b = ctypes.c_size_t()
        self._c(lib().hgpm_ensemble_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def generate_ic(self, box, ipk_header, ipk_values, delta, dot_delta, fscal, deltaT):
        v = np.ascontiguousarray(ipk_values, np.float64)
        self._c(lib().hgpm_ensemble_generate_ic(self.h, int(box), _ptr(v), v.size, float(ipk_header[3]),
                                                float(ipk_header[1]), float(delta), float(dot_delta), float(fscal),
                                                float(deltaT)))

    def set_particles(self, box, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
        vel = np.ascontiguousarray(vel, np.float32).reshape(-1)
        n = self.params[box].np ** 3
        assert pos.size == 3 * n and vel.size == 3 * n
        self._c(lib().hgpm_ensemble_set_particles(self.h, int(box), _ptr(pos), _ptr(vel)))

    def get_particles(self, box):
        n = self.params[box].np ** 3
        pos = np.zeros((n, 3), np.float32)
        vel = np.zeros((n, 3), np.float32)
        self._c(lib().hgpm_ensemble_get_particles(self.h, int(box), _ptr(pos), _ptr(vel)))
        return pos, vel

    def power_spectrum(self, box, fold=0):
        nb = self.params[box].pk_bins
        k, pk, cnt = np.zeros(nb), np.zeros(nb), np.zeros(nb, np.int64)
        self._c(lib().hgpm_ensemble_power_spectrum(self.h, int(box), int(fold), _ptr(k), _ptr(pk), _ptr(cnt)))
        return k, pk, cnt

    def step(self, coeffs):
        """coeffs: one StepCoeffs per box (equal within a batch)."""
        arr = (StepCoeffs * len(self.params))(*coeffs)
        self._c(lib().hgpm_ensemble_step(self.h, arr))

    def run(self, coeffs_per_step):
        """Several steps as one HIP graph: coeffs_per_step[s] is the per-box list
        step() takes for step s."""
        steps = [list(c) for c in coeffs_per_step]
        nb = len(self.params)
        assert all(len(c) == nb for c in steps)
        arr = (StepCoeffs * (nb * len(steps)))(*[q for c in steps for q in c])
        self._c(lib().hgpm_ensemble_run(self.h, len(steps), arr))

    def synchronize(self):
        self._c(lib().hgpm_ensemble_synchronize(self.h))

    def autotune(self, reps=3):
        self._c(lib().hgpm_ensemble_autotune(self.h, int(reps)))


# ------------------------------------------------- reference-style API ----
def UpdatePositions(params, mem, ts, frac):
    """serial::UpdatePositions (pm_actions.cpp:287-314)."""
    mem.update_positions(ts.drift_prefactor(frac))


def CIC(params, mem):
    """serial::CIC (pm_actions.cpp:221-258)."""
    mem.cic()


def SolveGradient(params, mem):
    """serial::SolveGradient (solver.cpp:84-98, 145-170)."""
    mem.solve_gradient()


def UpdateVelocities(params, mem, ts):
    """serial::UpdateVelocities (pm_actions.cpp:260-285): v += w dt fscal F."""
    mem.update_velocities(ts.deltaT, ts.fscal)


def GetPowerSpectrum(params, mem, fname):
    """serial::GetPowerSpectrum (power.cu:152-246): writes <fname>.fold.<f>."""
    mem.write_power_spectrum(fname)


def writeOutput(params, mem, fname):
    """serial::writeOutput (io.cu:8-44)."""
    mem.write_output(fname)


def PyAnalysis(params, mem, ts, analysis, step):
    """serial::PyAnalysis (src/analysis/cuda/py_analysis.cu:3-26): with DO_ANALYSIS set
    and `step` listed in ANALYSIS_STEPS, calls analysis(particles, n, 7, step, a, z) with
    the (n, 7) float32 array of mem.analysis_particles(); `analysis` stands for the
    analysisStep of the reference's ANALYSIS_PY module.  Returns whether it was called."""
    if not (params.do_analysis and params.analysis[step]):
        return False
    parts = mem.analysis_particles()
    analysis(parts, parts.shape[0], 7, step, ts.aa, ts.z)
    return True


def GenerateDisplacementIC(mem, params, ts, ipk_header, ipk_values):
    """serial::GenerateDisplacementIC (initializer.cpp:65-166) with the FP64 growth
    factor: delta at z_ini, dot_delta at a_ini - dt/2 (initializer.cpp:81-86)."""
    a_ini = 1.0 / (params.z_ini + 1.0)
    delta, _ = growth(params, a_ini)
    _, dot_delta = growth(params, a_ini - ts.deltaT / 2.0)
    mem.generate_ic(ipk_header, ipk_values, delta, dot_delta, ts.fscal, ts.deltaT)
    return delta, dot_delta


def make_params(ng, np_, rl, z_ini, z_fin=0.0, nsteps=10, seed=1, precision=FP32, pk_bins=32, pk_folds=0,
                omega_cdm=0.261, deut=0.0226576, hubble=0.68, ss8=0.81, ns=0.96, t_cmb=2.726,
                neff_massless=3.04, prefix="hgpm_run", ipk=""):
    """Params built in memory with read_params' float rounding and derived values."""
    f = lambda x: float(np.float32(x))
    p = Params()
    p.ng, p.np, p.rl, p.seed = int(ng), int(np_), f(rl), int(seed)
    p.z_ini, p.z_fin, p.nsteps, p.last_step = f(z_ini), f(z_fin), int(nsteps), int(nsteps)
    p.omega_cdm, p.deut, p.omega_nu, p.hubble = f(omega_cdm), f(deut), 0.0, f(hubble)
    p.ss8, p.ns, p.w_de, p.wa_de, p.t_cmb = f(ss8), f(ns), f(-1.0), 0.0, f(t_cmb)
    p.neff_massless, p.neff_massive = f(neff_massless), 0.0
    p.omega_baryon = p.deut / p.hubble / p.hubble
    p.omega_cb = p.omega_cdm + p.omega_baryon
    p.omega_matter = p.omega_cb + p.omega_nu
    p.omega_radiation = 2.471e-5 * (p.t_cmb / float(np.float32(2.725))) ** 4.0 / p.hubble ** 2.0
    p.f_nu_massless = p.neff_massless * 7.0 / 8.0 * (4.0 / 11.0) ** (4.0 / 3.0)
    p.f_nu_massive = 0.0
    p.pk_bins, p.pk_folds, p.block_size, p.ol = int(pk_bins), int(pk_folds), 64, 1.0
    p.precision = int(precision)
    p.prefix = prefix.encode()
    p.ipk = ipk.encode()
    return p
