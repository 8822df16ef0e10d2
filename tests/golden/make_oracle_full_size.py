"""Digests of the oracle's state after long runs, for bitwise GPU checks over whole meshes.

The oracle (oracle/, the CPU restatement of the reference) runs the two single-mesh
workloads at their full sizes -- the 3D bi-material tube 256x128x128 for 100 iterations
and the 3D Sedov blast 192^3 for 200 iterations -- and two runs of the served stream whose
cells empty out, to their final time, with the correctly rounded cube root that the GPU
library uses.  The fixture keeps, per parity field, the SHA-256 of the
float64 array (global cell / node order), its sum, and the x-profile at j = k = 0, plus
time, time step and iteration as exact hex floats.  A GPU run that reproduces every digest
is bitwise identical to the oracle over the whole mesh.

    python tests/golden/make_oracle_full_size.py [bisod3d] [sedov3d] [served114] [served162] [bisod3d_csts] [rider512]
                                                 [batch218] [batch144]
    (minutes on 8 threads)
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mahyco_amd import cases  # noqa: E402
from tests.helpers import summarize  # noqa: E402

def _served(i):
    """Request i of the seeded served stream (stiffened-gas Riemann runs to t = 0.1)."""
    from mahyco_amd import serve
    return serve.serve_requests(256, seed=20261202, rate=8.0)[i][1]


WORKLOADS = {
    "bisod3d": (lambda: cases.config2_bisod3d(256, 128, 128), 100),
    "sedov3d": (lambda: cases.config3_sedov3d(192, 200), 200),
    # two served runs whose cells empty out (total mass below the remap threshold squared):
    # the mass fractions then keep their values across the membership change
    "served114": (lambda: _served(114), 400),
    "served162": (lambda: _served(162), 400),
    # the CSTS scheme on the full bi-material tube, and RIDER advection (reversed
    # deformation, about 2e4 mixed cells) on a 512 x 512 mesh
    "bisod3d_csts": (lambda: _csts(cases.config2_bisod3d(256, 128, 128)), 100),
    "rider512": (lambda: _rider512(), 300),
    # the smallest (64 x 64) and the largest (512 x 512) member of the 256-run batched ensemble
    # (seed 20261202), 300 iterations each
    "batch218": (lambda: _batch(218), 300),
    "batch144": (lambda: _batch(144), 300),
}


def _batch(i):
    from mahyco_amd import serve
    return serve.batch_cases(256, seed=20261202)[i]


def _csts(c):
    c.schema_csts = True
    c.name += "_csts"
    return c


def _rider512():
    c = cases.rider("rider_deformation_reverse_512", 38, 1., 1.)
    c.n = (512, 512, 1)
    return c


def main(names):
    from oracle.pyoracle import OracleCase
    for name in names:
        make, cycles = WORKLOADS[name]
        case = make()
        t0 = time.time()
        o = OracleCase(case, pow_mode=1, threads=os.cpu_count() or 1)
        st = o.run(cycles)
        res = {"workload": case.name, "cycles": cycles, "status": int(st), "iteration": o.iteration,
               "time": float(o.time).hex(), "deltat": float(o.deltat).hex(), "fields": summarize(o, case)}
        o.close()
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"oracle_{name}_full.json")
        with open(path, "w") as fh:
            json.dump(res, fh, indent=0)
        print(f"{name}: {cycles} iterations, status {st}, {time.time() - t0:.0f} s -> {path}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or list(WORKLOADS))
