"""Host-side logic of libhgpm.so on the CPU: exported symbols, params/.ipk formats,
timestepper, growth, the Eisenstein-Hu .ipk writer and the error paths."""
import ctypes
import json
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

f32 = lambda x: float(np.float32(x))


def test_library_exports_every_declared_symbol(hg):
    L = hg.lib()
    names = hg.exported_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # nm agrees (dynamic symbol table)
    out = subprocess.run(["nm", "-D", "--defined-only", hg.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert f" T {n}\n" in out or f" T {n}" in out, n
    assert hg.version().startswith("hgpm")


def test_read_reference_params_fixture(hg):
    p = hg.read_params(os.path.join(GOLDEN, "test_run_params"))
    assert (p.ng, p.np, p.nsteps, p.last_step, p.seed, p.pk_folds, p.block_size) == (32, 32, 625, 625, 21082001, 1, 64)
    # FP64 fields carry float rounding (params.cpp:109-166)
    assert p.omega_cdm == f32(0.22) and p.hubble == f32(0.71) and p.rl == 8.0 and p.z_ini == 200.0
    assert p.omega_baryon == p.deut / p.hubble / p.hubble
    assert p.omega_cb == p.omega_cdm + p.omega_baryon and p.omega_matter == p.omega_cb + p.omega_nu
    assert abs(p.omega_radiation - 2.471e-5 * (f32(2.726) / f32(2.725)) ** 4 / f32(0.71) ** 2) < 1e-18
    assert abs(p.f_nu_massless - f32(3.04) * 7 / 8 * (4 / 11) ** (4 / 3)) < 1e-15
    dumps = [i for i in range(1000) if p.pk_dump[i]]
    assert dumps == [0, 1, 3, 5, 10, 20, 30, 40, 59, 66, 74, 86, 101, 122, 153, 175, 205, 247, 310, 415, 610, 624]
    assert p.prefix == b"runs/test3" and p.pk_bins == 221 and p.precision == hg.FP32
    assert not any(p.particle_dump) and p.dump_init == 0
    # analysis hook keys (params.cpp:195-214): "ANALYSIS_STEPS all", ANALYSIS_PY set,
    # ANALYSIS_DIR and DO_ANALYSIS commented out
    assert all(p.analysis) and p.analysis_py == b"/home/hqureshi/HACC_GPM/example_analysis/basic"
    assert p.analysis_dir == b"" and p.do_analysis == 0


def test_read_params_analysis_keys(hg, tmp_path):
    f = tmp_path / "a.params"
    f.write_text("NG 16\nNP 16\nANALYSIS_STEPS 0,7,999\nANALYSIS_DIR /tmp/x\nANALYSIS_PY mod.py\nDO_ANALYSIS\n")
    p = hg.read_params(str(f))
    assert [i for i in range(1000) if p.analysis[i]] == [0, 7, 999]
    assert p.analysis_dir == b"/tmp/x" and p.analysis_py == b"mod.py" and p.do_analysis == 1


def test_save_params_roundtrip(hg, tmp_path):
    p = hg.read_params(os.path.join(GOLDEN, "test_run_params"))
    p.prefix = str(tmp_path / "run").encode()
    hg.save_params(p)
    txt = (tmp_path / "run.params").read_text().splitlines()
    kv = dict(l.split(" ", 1) for l in txt)
    # the keys postprocessing/run.py reads back
    for k in ("NG", "NP", "RL", "Z_IN", "Z_FIN", "N_STEPS", "PK_FOLDS", "OUTPUT_BASE_NAME"):
        assert k in kv
    assert int(kv["NG"]) == 32 and float(kv["Z_IN"]) == 200 and int(kv["N_STEPS"]) == 625


def test_params_errors(hg, tmp_path):
    with pytest.raises(hg.HgpmError) as e:
        hg.read_params(tmp_path / "missing")
    assert e.value.code == hg.ERR_IO
    f = tmp_path / "nonp"
    f.write_text("NG 32\nRL 8\n")
    with pytest.raises(hg.HgpmError) as e:
        hg.read_params(f)
    assert e.value.code == hg.ERR_ARG
    f.write_text("NG 32\nNP 32\nPK_DUMP 1,2000\n")
    with pytest.raises(hg.HgpmError) as e:
        hg.read_params(f)
    assert e.value.code == hg.ERR_ARG and "out of range" in str(e.value)
    f.write_text("NG 32\nNP 32\nPRECISION FP16\n")
    with pytest.raises(hg.HgpmError):
        hg.read_params(f)
    f.write_text("NG 32\nNP 32\nUNKNOWN_KEY 5\n# comment\nPK_BINS 32\nPRECISION FP64\nPARTICLE_DUMP all\n")
    p = hg.read_params(f)
    assert p.pk_bins == 32 and p.precision == hg.FP64 and all(p.particle_dump)


def test_read_reference_ipk_fixture(hg, tmp_path):
    hdr, vals = hg.read_ipk(os.path.join(GOLDEN, "test_run_ipk.ipk"))
    assert hdr[0] == 0 and hdr[1] == 1000 and hdr[2] == 20000 and abs(hdr[3] - 1000 / 19999) < 1e-12
    assert hdr[15] == 200 and vals.size == 20000 and vals[0] == 0 and abs(vals[1] - 0.5576035) < 1e-6
    short = tmp_path / "short.ipk"
    short.write_bytes(np.zeros(20 + 5).tobytes()[:100])
    with pytest.raises(hg.HgpmError) as e:
        hg.read_ipk(short)
    assert e.value.code == hg.ERR_IO
    with pytest.raises(hg.HgpmError):
        hg.read_ipk(tmp_path / "nope.ipk")


def test_timestepper_closed_form(hg):
    p = hg.make_params(64, 64, 128.0, 200.0, 0.0, 60)
    # adot for LCDM + radiation: sqrt((Ocb + (1+fnu) Or/a + OL a^3)/a) with OL = 1 - Om - (1+fnu) Or
    a = 0.37
    orad = p.omega_radiation * (1 + p.f_nu_massless)
    ol = 1 - p.omega_matter - orad
    expect = math.sqrt((p.omega_cb + orad / a + ol * a ** 3) / a)
    assert abs(hg.get_adot(p, a) - expect) < 1e-15
    fs = hg.get_fscal(a, expect, p.omega_cb)
    assert abs(fs - 1.5 * p.omega_cb * float(np.float32(1 / expect)) / a) < 1e-15
    ts = hg.Timestepper(p)
    assert ts.deltaT == ((1 / (0.0 + 1)) - (1 / (200.0 + 1))) / 60
    ts.setInitialZ(200.0)
    a0 = ts.aa
    ts.reverseHalfStep()
    ts.advanceHalfStep()
    for _ in range(120):
        ts.advanceHalfStep()
    assert abs(ts.aa - (a0 + 60 * ts.deltaT)) < 1e-12 and abs(ts.aa - 1.0) < 1e-12


def test_schedule_matches_timestepper(hg):
    p = hg.make_params(64, 64, 128.0, 200.0, 0.0, 60)
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    ts.advanceHalfStep()
    co = hg.schedule(p, 60, merge=False)
    for s in range(60):
        assert co[s].drift0 == ts.drift_prefactor(0.5) and co[s].skip_drift0 == 0
        ts.advanceHalfStep()
        assert co[s].kick_dt == ts.deltaT and co[s].kick_fscal == ts.fscal
        ts.advanceHalfStep()
        assert co[s].drift1 == ts.drift_prefactor(0.5)
    out = np.zeros(60, np.uint8)
    out[10] = 1
    m = hg.schedule(p, 60, merge=True, output_after=out)
    assert m[0].skip_drift0 == 0 and m[1].skip_drift0 == 1 and m[0].drift1 == 2 * co[0].drift1
    assert m[10].drift1 == co[10].drift1 and m[11].skip_drift0 == 0 and m[11].drift0 == co[11].drift0
    assert m[59].drift1 == co[59].drift1


def test_growth_and_eh_against_golden(hg, orc, tmp_path):
    g = json.load(open(os.path.join(GOLDEN, "cosmo_golden.json")))
    p = hg.make_params(64, 64, 128.0, 50.0)
    assert abs(p.omega_matter - g["cosmology"]["omega_matter"]) < 1e-15
    for e in g["growth"]:
        D, dD = hg.growth(p, e["a"])
        assert abs(D / e["D"] - 1) < 1e-9 and abs(dD / e["dDda"] - 1) < 1e-6
        Do, dDo = orc.growth_f64(e["a"], p.omega_matter, 1 - p.omega_matter)
        assert abs(D / Do - 1) < 1e-10 and abs(dD / dDo - 1) < 1e-8
    fn = tmp_path / "eh.ipk"
    hg.write_ipk_eh(p, fn, 100.0, 100001)
    hdr, vals = hg.read_ipk(fn)
    assert hdr[2] == 100001 and abs(hdr[3] - 0.001) < 1e-15 and hdr[15] == f32(50.0) and hdr[4] == p.omega_cdm
    ks = np.linspace(0, 100, 100001)
    for k, P in zip(g["pk_z50"]["k"], g["pk_z50"]["P"]):
        i = int(round(k / 0.001))
        assert abs(ks[i] - k) < 1e-12
        assert abs(vals[i] / P - 1) < 2e-6, (k, vals[i], P)


def test_create_without_gpu_fails_loudly(hg):
    # the library's own device query: no second HIP runtime (torch) in this process
    if hg.device_count() > 0:
        pytest.skip("a GPU is visible")
    p = hg.make_params(32, 32, 64.0, 50.0)
    with pytest.raises(hg.HgpmError) as e:
        hg.MemoryManager(p)
    assert e.value.code == hg.ERR_HIP


def test_create_rejects_bad_sizes(hg):
    for ng, np_ in ((33, 33), (64, 128), (2048, 2048)):
        p = hg.make_params(ng, np_, 64.0, 50.0)
        with pytest.raises(hg.HgpmError) as e:
            hg.MemoryManager(p)
        assert e.value.code in (hg.ERR_ARG, hg.ERR_UNSUPPORTED)


def test_driver_refuses_multi_rank_and_missing_file(hg, tmp_path):
    r = subprocess.run([hg.DRIVER_PATH, str(tmp_path / "none.params")], capture_output=True, text=True)
    assert r.returncode == 1 and "not found" in r.stderr
    r = subprocess.run([hg.DRIVER_PATH, "x.params", "+n4"], capture_output=True, text=True)
    assert r.returncode == 1 and "one GPU" in r.stderr
