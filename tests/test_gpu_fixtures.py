"""The HIP path against the reference's own non-regression outputs (GPU tests), with the
same bars as the oracle: final time to 1e-13, every time-history time stamp to 1e-13 and
EnvSummation value to 1e-12 (relative; 2.5e-11 for the 216000-cell Sedov run), final
EnSight fields within printed precision."""
import os

import numpy as np
import pytest

from mahyco_amd import types as T
from tests.test_oracle_fixtures import ENSIGHT, FIXTURES, GOLDEN, TH, TH_RTOL, _printed

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(FIXTURES))
def test_gpu_matches_reference_fixture(name):
    from mahyco_amd.lib import GpuCase
    print(f"[gpu fixture] {name}", flush=True)
    ref = np.load(os.path.join(GOLDEN, name + ".npz"))
    case = FIXTURES[name]()
    g = GpuCase(case)
    g.history_enable(10, 1024)
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
                np.testing.assert_allclose(th[:, 1 + 5 * k + vi], ref[key], rtol=TH_RTOL.get(name, 1e-12), atol=0,
                                           err_msg=key)
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
