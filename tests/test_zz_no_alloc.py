"""hgpm_generate_ic_noalloc (the initial conditions with the .ipk table staged in the
device's fixed array, no device allocation) against hgpm_generate_ic on the same inputs:
bit-identical particles, on two calls in a row, the second with a smaller context and a
smaller table; FP64 contexts too; a table over the fixed size is refused."""
import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(hg, p, k_max, bins):
    ts = hg.Timestepper(p)
    ts.setInitialZ(p.z_ini)
    ts.reverseHalfStep()
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "eh.ipk")
        hg.write_ipk_eh(p, fn, k_max, bins)
        hdr, vals = hg.read_ipk(fn)
    a = 1.0 / (1.0 + p.z_ini)
    delta, _ = hg.growth(p, a)
    _, dot_delta = hg.growth(p, a - ts.deltaT / 2.0)
    return hdr, vals, delta, dot_delta, ts.fscal, ts.deltaT


def test_generate_ic_noalloc_matches_generate_ic(hg):
    # (ng, np, precision, .ipk bins): the second call is smaller in particles and table
    cases = [(64, 64, hg.FP32, 20001), (32, 32, hg.FP32, 5001), (32, 32, hg.FP64, 5001)]
    for ng, np_, prec, bins in cases:
        print(f"[case] generate_ic_noalloc ng={ng} np={np_} prec={prec} bins={bins}", flush=True)
        p = hg.make_params(ng, np_, 2.0 * ng, 50.0, 0.0, 100, seed=ng + bins, precision=prec)
        args = _inputs(hg, p, 20.0, bins)
        with hg.MemoryManager(p) as ref, hg.MemoryManager(p) as m:
            ref.generate_ic(*args)
            m.generate_ic_noalloc(*args)
            if prec == hg.FP32:
                for a, b in zip(ref.get_particles_split(), m.get_particles_split()):
                    assert np.array_equal(a, b)
            else:
                for a, b in zip(ref.get_particles_f64(), m.get_particles_f64()):
                    assert np.array_equal(a, b)
    # a table over the fixed size is refused before anything runs
    p = hg.make_params(32, 32, 64.0, 50.0, 0.0, 100)
    hdr, vals, *rest = _inputs(hg, p, 20.0, 5001)
    big = np.zeros(hg.IC_STATIC_BINS + 1)
    big[: vals.size] = vals
    with hg.MemoryManager(p) as m:
        with pytest.raises(hg.HgpmError) as e:
            m.generate_ic_noalloc(hdr, big, *rest)
        assert e.value.code == hg.ERR_ARG and "staging array" in str(e.value)
