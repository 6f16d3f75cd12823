"""Pin the CPU oracle against the reference's golden restart files.

Each case is run exactly as the reference's regression test does
(runs/vquad.tst, runs/vhex.tst) and compared with ``dgm_diff`` semantics
(src/Domain.cpp:1000-1071): every field must pass the absolute tolerance and,
where the reference norm exceeds repsM = sqrt(eps), the relative tolerance.
"""
import numpy as np
import pytest

from dgm_amd import formats
from fixture_runs import GOLDEN, run_oracle_case
from oracle_lib import dgm_diff

REPS_M = np.sqrt(np.finfo(float).eps)


def _check(root, aeps, reps):
    o, Uh = run_oracle_case(root)
    ref = formats.read_restart(f"{GOLDEN}/{root}.ref", o.ne)
    rn, dn = dgm_diff(o, ref.data, Uh)
    assert np.all(dn <= aeps), f"{root}: absolute diff {dn}"
    mask = rn > REPS_M
    rel = np.where(mask, dn / np.where(mask, rn, 1.0), 0.0)
    assert np.all(rel <= reps), f"{root}: relative diff {rel}"
    return dn, rel


def test_vquad_2d_ns_br1_tvd():
    # runs/vquad.tst:23-24 aeps = reps = 5e-12
    _check("vquad", 5e-12, 5e-12)


def test_vhex_3d_euler_lf_tvd():
    # runs/vhex.tst:23-24 aeps = 5e-13, reps = 5e-10
    _check("vhex", 5e-13, 5e-10)


def test_restart_reader_consumes_whole_files():
    for root, ne in (("vquad", 4), ("vhex", 8), ("nvortex", 25)):
        r = formats.read_restart(f"{GOLDEN}/{root}.ref", ne)
        assert r.data.shape[1] == ne
        assert np.all(np.isfinite(r.data))
