"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host code, g++): a driver
built from the same sources (oracle/sanitize.mk) runs small cases of every remap and Lagrange
variant; no sanitizer report, and the same iteration, time and density sum as the oracle
library built without sanitizers."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from mahyco_amd import cases
from mahyco_amd import types as T
from oracle.pyoracle import OracleCase

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
DRIVER = os.path.join(ORACLE, "_build", "sanitize_driver")


def _small(c, n):
    c.n = n
    return c


def _case_bytes(case):
    mesh, envs, opt = case.mesh_c(), case.envs_c(), case.options_c()
    return (bytes(mesh) + C.c_int(len(case.envs)).value.to_bytes(4, "little") +
            b"".join(bytes(e) for e in envs) + bytes(opt))


CASES = [
    ("bisod3d_superbee", lambda: cases.bisod_x(24, 3, 3, limiter=T.MH_SUPERBEE), 25),
    ("bitriplepoint_vanleer_3env", lambda: _small(cases.bitriplepoint(), (35, 15, 1)), 20),
    ("order3", lambda: cases.ordre3(lambda: cases.bisod_x(24, 3, 3), "o3"), 20),
    ("pente_borne_debar2", lambda: cases.pente_borne(lambda: cases.bisod_x(24, 3, 3), "pb", 2, T.MH_SUPERBEE), 20),
    ("csts", lambda: _csts(cases.bisod_x(24, 3, 3)), 20),
    ("lagrange_newton", cases.lag2d_sod_x_newton, 20),
    ("rider_advection", lambda: _small(cases.rider("r", 34, .2), (24, 24, 1)), 20),
    ("sedov2d", lambda: _small(cases.sedov_2d(), (24, 24, 1)), 20),
]


def _csts(c):
    c.schema_csts = True
    return c


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-s", "-C", ORACLE, "-f", "sanitize.mk"], check=True)
    return DRIVER


@pytest.mark.parametrize("name,make,cycles", CASES, ids=[c[0] for c in CASES])
def test_oracle_clean_under_sanitizers(driver, tmp_path, name, make, cycles):
    case = make()
    path = tmp_path / f"{name}.case"
    path.write_bytes(_case_bytes(case))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([driver, str(path), str(cycles)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    kv = dict(x.split("=") for x in r.stdout.split())
    o = OracleCase(case, threads=1)
    st = o.run(cycles)
    assert int(kv["iteration"]) == o.iteration and int(kv["status"]) == st
    assert float(kv["time"]) == o.time
    assert float(kv["sum_rho"]) == float(np.cumsum(o.get(T.MH_F_DENSITY))[-1])  # summed in cell order
