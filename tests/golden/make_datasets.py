"""Copies the reference's non-regression datasets (NONREGRESSION/<case>/Donnees.arc, the
inputs of its own tests) that the dataset-reader tests read: every case with a fixture, and
three whose options this path refuses (a Mie-Grueneisen EOS, an ALE remap and a tabulated
EOS).

    python tests/golden/make_datasets.py /root/reference/NONREGRESSION
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

REFUSED = ["CAS_Compression_isotherme_ALU_monodimX", "CAS_2D_ALE_Sedov_euler", "CAS_2D_choc_plaque_crenau"]


def main(src):
    from tests.test_oracle_fixtures import FIXTURES
    out = os.path.join(HERE, "datasets")
    os.makedirs(out, exist_ok=True)
    for name in sorted(FIXTURES) + REFUSED:
        shutil.copyfile(os.path.join(src, name, "Donnees.arc"), os.path.join(out, name + ".arc"))
    print(f"{len(FIXTURES) + len(REFUSED)} datasets in {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/NONREGRESSION")
