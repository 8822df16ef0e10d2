"""Packs the reference's own output files of three non-regression cases as fixtures.

NONREGRESSION/<case>/output holds what the reference wrote for the case: the EnSight Gold
post-processing files (output/depouillement) and the EnvSummation time histories
(output/courbes/gnuplot), which NONREGRESSION/launch_test.sh:74-96 compares with diff -r.
They are copied unchanged into tests/golden/outputs/<case>.tar.gz.

    python tests/golden/make_output_fixtures.py /root/reference/NONREGRESSION
"""
import os
import sys
import tarfile

CASES = ["CAS_BiSodCaseX", "CAS_2D_Lag_SodCaseX", "CAS_2D_Sedov"]


def main(src):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "outputs")
    os.makedirs(out, exist_ok=True)
    for c in CASES:
        root = os.path.join(src, c, "output")
        path = os.path.join(out, c + ".tar.gz")
        with tarfile.open(path, "w:gz") as t:
            for sub in ("depouillement", "courbes"):
                t.add(os.path.join(root, sub), arcname=sub)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/NONREGRESSION")
