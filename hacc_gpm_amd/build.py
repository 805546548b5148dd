"""Builds the engine: hacc_gpm_amd/libhgpm.so (HIP, gfx950) and hacc_gpm_amd/bin/haccgpm.

Every HIP source is compiled with hipcc for gfx950 only, with packed-FP32 ops
disabled; the FFT unit is compiled once per supported transform length.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhgpm.so")
BIN = os.path.join(HERE, "bin", "haccgpm")
ROOT = os.path.dirname(HERE)

FFT_SIZES = [32, 64, 96, 128, 192, 256, 384, 512, 768, 1024]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def _jobs():
    jobs = []
    for n in FFT_SIZES:
        jobs.append((os.path.join(CSRC, "k_fft.hip"), os.path.join(OBJ, f"k_fft_{n}.o"), [f"-DHGPM_FFT_N={n}"]))
    # the alternative radix order of the 1024-point transforms
    jobs.append((os.path.join(CSRC, "k_fft.hip"), os.path.join(OBJ, "k_fft_1024_alt.o"),
                 ["-DHGPM_FFT_N=1024", "-DHGPM_FFT_PLAN_ALT"]))
    for name in ["k_particles.hip", "k_power.hip", "k_ic.hip", "k_fp64.hip", "capi.cpp", "host.cpp", "ensemble.cpp"]:
        src = os.path.join(CSRC, name)
        if os.path.exists(src):
            base = os.path.splitext(name)[0]
            lang = ["-x", "hip"] if name.endswith(".hip") else []
            jobs.append((src, os.path.join(OBJ, base + ".o"), lang))
    return jobs


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    todo = _jobs()
    nj = jobs or min(8, os.cpu_count() or 4)

    def one(j):
        src, obj, extra = j
        return _run([HIPCC] + HIP_FLAGS + extra + ["-c", src, "-o", obj])

    with ThreadPoolExecutor(nj) as ex:
        for out in ex.map(one, todo):
            if verbose and out.strip():
                print(out)
    objs = [o for _, o, _ in todo]
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", LIB] + objs)
    _run([HIPCC, "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
          os.path.join(CSRC, "driver_pm.cpp"), "-o", BIN, "-L" + HERE, "-lhgpm",
          "-Wl,-rpath,$ORIGIN/.."])
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
