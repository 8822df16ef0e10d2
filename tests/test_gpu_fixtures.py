"""The HIP path against the reference's own non-regression outputs (GPU tests), with the
same bars as the oracle: final time to 1e-13, every time-history time stamp to 1e-13 and
EnvSummation value to 1e-12 (relative; the same per-run exceptions for ill-conditioned
averages), final EnSight fields within printed precision (except for the chaotic run)."""
import os

import numpy as np
import pytest

from mahyco_amd import types as T
from tests.test_oracle_fixtures import CHAOTIC, ENSIGHT, FIXTURES, GOLDEN, RESTART_AFTER, TH, _printed, th_bar

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(FIXTURES))
def test_gpu_matches_reference_fixture(name, tmp_path):
    from mahyco_amd.lib import GpuCase
    print(f"[gpu fixture] {name}", flush=True)
    ref = np.load(os.path.join(GOLDEN, name + ".npz"))
    case = FIXTURES[name]()
    g = GpuCase(case)
    g.history_enable(10, 1024)
    if name in RESTART_AFTER:  # stop, dump, restore into a new case, continue as the reference
        g.run(RESTART_AFTER[name], graph=6)
        g.checkpoint_write(tmp_path / "dump")
        g.close()
        g = GpuCase(case)
        g.history_enable(10, 1024)
        g.checkpoint_read(tmp_path / "dump")
        g.continue_init()
    st = T.MH_OK
    while st == T.MH_OK:
        st = g.run(500, graph=6)
    assert st == T.MH_STOPPED, st
    ft = float(ref["final_time"])
    assert abs(g.time - ft) <= 1e-13 * ft, (g.time, ft)
    th = g.history()
    np.testing.assert_allclose(th[:, 0], ref["th_time"], rtol=1e-13, atol=0)
    for k in range(len(ref["env_names"])):
        for vi, v in enumerate(TH):
            key = f"th_{v}_{k}"
            if key in ref:
                rows, rtol = th_bar(name, v)
                np.testing.assert_allclose(th[:rows, 1 + 5 * k + vi], ref[key][:rows], rtol=rtol, atol=0,
                                           err_msg=key)
    if name in CHAOTIC:
        g.close()
        return
    for var, fid in ENSIGHT.items():
        r = ref["cell_" + var]
        d = np.abs(_printed(g.get(fid)) - r)
        assert np.all(d <= 5e-6 * np.abs(r) + 1e-12), (var, float(d.max()))
    for c in range(case.dim):
        if f"node_Velocity_{c}" not in ref:
            continue  # the run wrote no velocity
        r = ref[f"node_Velocity_{c}"]
        d = np.abs(_printed(g.get(T.MH_N_VELOCITY, -1, c)) - r)
        assert np.all(d <= 5e-6 * np.abs(r) + 1e-12), (c, float(d.max()))
    g.close()
