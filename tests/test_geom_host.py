"""The corner-vector forms of mahyco_amd/csrc/mh_geom.h on the host (g++, -ffp-contract=off as
the kernels are built): the component passes cqs_hex_comp<C> and their face-by-face form
cqs_hex_comp_fm<C> give cqs_hex's component C bit for bit (cqs_hex is the term-by-term
restatement of MahycoModule::computeCQs, MahycoModule.cc:1869-1932, pinned through the
oracle by the reference's fixtures).  Inputs: perturbed unit cubes at several scales and
offsets, twisted and collapsed cells, integer coordinates (many exact zeros and cancellations)
and signed zeros."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CUBE = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], dtype=np.float64)


def _cells(rng):
    out = []
    for scale, off in ((1., 0.), (1e-3, 0.5), (7.3, -3.), (1e4, 1e6), (1e-8, 1.)):
        for amp in (0., 1e-3, 0.2, 0.45):
            x = off + scale * (CUBE[None] + amp * rng.uniform(-1, 1, (200, 8, 3)))
            out.append(x)
    twist = CUBE[None].repeat(50, 0).copy()
    twist[:, 4:, :2] = twist[:, 4:, :2] @ np.array([[0, -1], [1, 0]]) + rng.uniform(-.3, .3, (50, 4, 2))
    out.append(twist)
    flat = CUBE[None].repeat(50, 0).copy()
    flat[:, :, rng.integers(0, 3)] *= 0.
    out.append(flat + rng.uniform(-1e-9, 1e-9, flat.shape) * (rng.random(flat.shape) < .5))
    out.append(rng.integers(-3, 4, (400, 8, 3)).astype(np.float64))
    z = rng.choice(np.array([0., -0., 1., -1., 0.5]), (400, 8, 3))
    out.append(z)
    return np.concatenate([c.reshape(-1, 8, 3) for c in out])


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("geom") / "geom_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(HERE, "host", "geom_check.cpp"),
                    "-o", exe], check=True)
    return exe


def test_component_forms_equal_cqs_hex_bitwise(checker):
    x = _cells(np.random.default_rng(20261203))
    words = x.reshape(-1).view(np.uint64)
    text = "\n".join(" ".join(f"{w:x}" for w in words[i:i + 24]) for i in range(0, len(words), 24))
    r = subprocess.run([checker], input=text, capture_output=True, text=True, check=True)
    cells, bad = map(int, r.stdout.split())
    assert cells == len(x)
    assert bad == 0
