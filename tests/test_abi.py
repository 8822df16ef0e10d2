"""The C-ABI library loads and exports every function include/*.h declares (no GPU calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mahyco_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from mahyco_amd import lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("HIP library not built")
    so = C.CDLL(L.LIB_PATH)
    names = _declared("mahyco.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_struct_layout_matches_header():
    """ctypes mirrors have the C sizes of the header structs (checked with the C compiler)."""
    import subprocess
    import tempfile
    from mahyco_amd import types as T
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "mahyco_types.h"
int main(void){ printf("%zu %zu %zu %zu %zu\n", sizeof(MhMesh), sizeof(MhEnvironment), sizeof(MhBoundaryCondition),
  sizeof(MhOptions), offsetof(MhOptions, sedov_deposit)); return 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "t.c")
        open(p, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, p], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    got = [C.sizeof(T.MhMesh), C.sizeof(T.MhEnvironment), C.sizeof(T.MhBoundaryCondition),
           C.sizeof(T.MhOptions), T.MhOptions.sedov_deposit.offset]
    assert [int(x) for x in out] == got


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from mahyco_amd import lib as L
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(RuntimeError):
        L.load()


def test_footprint_counts_variant_arrays():
    """Ensemble admission by HBM footprint (host-only call): every option that allocates more
    device arrays raises the estimate, by the arrays it adds."""
    from mahyco_amd import cases
    from mahyco_amd import lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("HIP library not built")
    base = cases.config2_bisod3d(32, 16, 16)
    nc, nn, m = 32 * 16 * 16, 33 * 17 * 17, 2
    f0 = L.footprint(base)

    def with_(**kw):
        c = cases.config2_bisod3d(32, 16, 16)
        for k, v in kw.items():
            setattr(c, k, v)
        return L.footprint(c)

    assert with_(schema_csts=True) - f0 == 8 * (3 * nn + 3 * nc) + nc
    assert with_(ordre_projection=3) - f0 == 8 * 3 * nc + nc  # cell widths, mixed flags
    assert with_(pente_borne=True) - f0 == 8 * (3 * nc + (3 * m + 1) * nc)  # widths, first-sweep phi
    assert L.footprint(cases.config2_bisod3d(64, 16, 16)) > f0
