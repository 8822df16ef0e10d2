"""On-disk outputs against the files the reference wrote for the same cases (GPU tests).

The reference's non-regression suite compares its EnSight post-processing and gnuplot time
histories with diff -r (NONREGRESSION/launch_test.sh:74-96).  Here the GPU run of three
cases writes both with mahyco_ensight_write / mahyco_history_write_gnuplot and they are
compared with the reference's own files (tests/golden/outputs, packed unchanged by
tests/golden/make_output_fixtures.py): the EnSight files byte for byte, the time histories
line for line in the same format, with the values at the bars of the fixture tests (time
stamps 1e-13, EnvSummation values 1e-12 relative: the last of their 17 printed digits
depend on summation order)."""
import os
import re
import tarfile

import numpy as np
import pytest

from tests.test_oracle_fixtures import FIXTURES, GOLDEN

pytestmark = pytest.mark.gpu

OUTPUTS = os.path.join(GOLDEN, "outputs")
CASES = ["CAS_BiSodCaseX", "CAS_2D_Lag_SodCaseX", "CAS_2D_Sedov"]
LINE = re.compile(r"^-?\d\.\d{16}E[+-]\d{2} -?\d\.\d{16}E[+-]\d{2}$")


def _files(root):
    out = {}
    for d, _, fs in os.walk(root):
        for f in fs:
            p = os.path.join(d, f)
            out[os.path.relpath(p, root)] = p
    return out


@pytest.mark.parametrize("name", CASES)
def test_outputs_match_reference_files(name, tmp_path):
    from mahyco_amd import types as T
    from mahyco_amd.lib import GpuCase
    print(f"[outputs] {name}", flush=True)
    ref = tmp_path / "ref"
    with tarfile.open(os.path.join(OUTPUTS, name + ".tar.gz")) as t:
        t.extractall(ref)
    case_text = (ref / "depouillement" / "ensight.case").read_text()
    variables = re.findall(r"^(?:scalar|vector) per \w+:\s+1\s+(\w+) ", case_text, re.M)
    env_names = [str(e) for e in np.load(os.path.join(GOLDEN, name + ".npz"))["env_names"]]
    g = GpuCase(FIXTURES[name]())
    g.history_enable(10, 4096)
    st = T.MH_OK
    while st == T.MH_OK:
        st = g.run(500, graph=6)
    assert st == T.MH_STOPPED
    out = tmp_path / "out"
    g.ensight_write(out / "depouillement", variables)
    g.history_write_gnuplot(out / "courbes" / "gnuplot", env_names)
    g.close()
    _compare(out, ref)


@pytest.mark.parametrize("name", CASES)
def test_dataset_run_writes_the_reference_files(name, tmp_path):
    """The dataset itself (tests/golden/datasets, the reference's Donnees.arc) through the
    reader and runner (python -m mahyco_amd.arc): the same files as the reference's run."""
    from mahyco_amd import arc
    from mahyco_amd import types as T
    print(f"[outputs] dataset {name}", flush=True)
    ref = tmp_path / "ref"
    with tarfile.open(os.path.join(OUTPUTS, name + ".tar.gz")) as t:
        t.extractall(ref)
    st, it, tm = arc.run_arc(os.path.join(GOLDEN, "datasets", name + ".arc"), str(tmp_path / "out"))
    assert st == T.MH_STOPPED
    _compare(tmp_path / "out", ref)


def _compare(out, ref):
    mine, theirs = _files(out), _files(ref)
    assert sorted(mine) == sorted(theirs)
    for rel in sorted(theirs):
        a = open(mine[rel]).read().splitlines()
        b = open(theirs[rel]).read().splitlines()
        assert len(a) == len(b), rel
        if rel.startswith("depouillement"):
            bad = [i for i, (x, y) in enumerate(zip(a, b)) if x != y]
            assert not bad, (rel, len(bad), [(a[i], b[i]) for i in bad[:4]])
            assert open(mine[rel], "rb").read() == open(theirs[rel], "rb").read(), rel
        else:
            assert all(LINE.match(x) for x in a), rel
            va = np.array([[float(v) for v in x.split()] for x in a])
            vb = np.array([[float(v) for v in x.split()] for x in b])
            np.testing.assert_allclose(va[:, 0], vb[:, 0], rtol=1e-13, atol=0, err_msg=rel)
            np.testing.assert_allclose(va[:, 1], vb[:, 1], rtol=1e-12, atol=0, err_msg=rel)


def _sedov3d_csts():
    from mahyco_amd import cases
    c = cases.config3_sedov3d(16)
    c.schema_csts = True
    return c


@pytest.mark.parametrize("make", [lambda: FIXTURES["CAS_BiSodCaseX"](), lambda: FIXTURES["CAS_bitriplepoint_minmod"](),
                                  lambda: FIXTURES["CAS_2D_Lag_SodCaseX"](), _sedov3d_csts])
def test_checkpoint_restart_is_bitwise(make, tmp_path):
    """Run 23 iterations, checkpoint, continue 31 in a fresh case restored from the file:
    every field, the time, the time step and the time history equal an uninterrupted run of
    54 iterations (hipGraph replays on both sides; 23 is not a multiple of the graph period)."""
    from mahyco_amd import types as T
    from mahyco_amd.lib import GpuCase, MahycoError
    from tests.helpers import bitwise_equal, fields
    case = make()
    print(f"[checkpoint] {case.name}", flush=True)
    a = GpuCase(case)
    a.history_enable(10, 64)
    a.run(54, graph=6)
    b = GpuCase(case)
    b.history_enable(10, 64)
    b.run(23, graph=6)
    path = tmp_path / "case.ckpt"
    b.checkpoint_write(path)
    b.close()
    c = GpuCase(case)
    c.history_enable(10, 64)
    c.checkpoint_read(path)
    assert c.iteration == 23
    c.run(31, graph=6)
    assert c.iteration == a.iteration and c.time == a.time and c.deltat == a.deltat
    fa, fc = fields(a, len(case.envs), case.dim), fields(c, len(case.envs), case.dim)
    for k in fa:
        assert bitwise_equal(fa[k], fc[k]), k
    assert bitwise_equal(a.history(), c.history())
    # a checkpoint of another case is refused and leaves the case as it was
    other = FIXTURES["CAS_2D_Sedov"]() if case.dim == 3 else FIXTURES["CAS_BiSodCaseX"]()
    d = GpuCase(other)
    with pytest.raises(MahycoError):
        d.checkpoint_read(path)
    assert d.iteration == 0
    for x in (a, c, d):
        x.close()


@pytest.mark.parametrize("nenv", [1, 2])
def test_continue_init_matches_oracle(nenv, tmp_path):
    """mahyco_checkpoint_read + mahyco_continue_init continue as the reference's restart does
    (hydroContinueInit, MahycoModule.cc:288: 3m+5 projected slots): bitwise equal to the
    oracle continued the same way; with one environment the pseudo slot is still projected
    and the run equals an uninterrupted one."""
    import numpy as np
    from mahyco_amd import cases
    from mahyco_amd.lib import GpuCase
    from oracle.pyoracle import OracleCase
    from tests.helpers import bitwise_equal, fields
    if nenv == 1:
        case = cases.config3_sedov3d(12)
    else:
        case = cases.riemann3d(12, np.random.default_rng(32), stiffened=True, t_end=1e30, name="restart2")
    print(f"[continue] m={len(case.envs)}", flush=True)
    g = GpuCase(case)
    g.run(7, graph=6)
    g.checkpoint_write(tmp_path / "dump")
    g.close()
    g = GpuCase(case)
    g.checkpoint_read(tmp_path / "dump")
    g.continue_init()
    g.run(40, graph=6)
    o = OracleCase(case, pow_mode=1)
    o.run(7)
    o.continue_after_restart()
    o.run(40)
    assert g.iteration == o.iteration and g.time == o.time
    fg, fo = fields(g, len(case.envs), 3), fields(o, len(case.envs), 3)
    for k in fo:
        assert bitwise_equal(fg[k], fo[k]), k
    if nenv == 1:
        u = OracleCase(case, pow_mode=1)
        u.run(47)
        fu = fields(u, 1, 3)
        for k in fo:
            assert bitwise_equal(fu[k], fo[k]), k
    g.close()
