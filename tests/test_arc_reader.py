"""The Arcane dataset reader (mahyco_amd/arc.py) on the reference's own datasets (CPU).

Every non-regression dataset with a fixture (tests/golden/datasets, copied unchanged by
tests/golden/make_datasets.py) reads into the same mesh, environment and option structures
as the case written out by hand for the fixture tests, which the fixtures pin; datasets that
ask for what this path does not run are refused with the option named."""
import ctypes as C
import os

import numpy as np
import pytest

from mahyco_amd import types as T
from mahyco_amd.arc import ArcError, ArcUnsupported, read_arc
from tests.test_oracle_fixtures import FIXTURES

DATASETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "datasets")

# The triple point declares pente-borne-mixte, a debar fix and a pure-neighbourhood limiter with
# pente-borne off; they switch nothing (RemapADIService.cc:203-209 needs limiteur-pure-id 1 for the
# mixte variant, the debar fix belongs to pente-borne), so its hand-written case leaves them at 0.
INERT = {"CAS_bitriplepoint_minmod": {"opt.pente_borne_debar_fix", "opt.pente_borne_mixte",
                                      "opt.projection_limiteur_pure_id"}}


def _flat(s, prefix):
    out = {}
    for f, _ in s._fields_:
        v = getattr(s, f)
        if hasattr(v, "_fields_"):
            out.update(_flat(v, f"{prefix}.{f}"))
        elif isinstance(v, C.Array):
            for i in range(len(v)):
                x = v[i]
                out.update(_flat(x, f"{prefix}.{f}[{i}]") if hasattr(x, "_fields_") else {f"{prefix}.{f}[{i}]": x})
        else:
            out[f"{prefix}.{f}"] = v
    return out


def _structs(case):
    d = _flat(case.mesh_c(), "mesh")
    d.update(_flat(case.options_c(), "opt"))
    for i, e in enumerate(case.envs_c()):
        d.update(_flat(e, f"env{i}"))
    return d


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_dataset_reads_as_the_fixture_case(name):
    got = read_arc(os.path.join(DATASETS, name + ".arc"))
    ref = FIXTURES[name]()
    assert len(got.envs) == len(ref.envs)
    a, b = _structs(got), _structs(ref)
    diff = {k for k in a if a[k] != b[k]}
    assert diff == INERT.get(name, set()), {k: (a[k], b[k]) for k in diff}


def test_inert_options_change_nothing():
    from oracle.pyoracle import OracleCase
    from tests.helpers import fields
    name = "CAS_bitriplepoint_minmod"
    a, b = FIXTURES[name](), read_arc(os.path.join(DATASETS, name + ".arc"))
    oa, ob = OracleCase(a, threads=2), OracleCase(b, threads=2)
    assert oa.run(60) == ob.run(60) == T.MH_OK
    fa, fb = fields(oa, 3, 2), fields(ob, 3, 2)
    assert oa.time == ob.time and all(np.array_equal(fa[k], fb[k]) for k in fa)


def test_outputs_and_names():
    c = read_arc(os.path.join(DATASETS, "CAS_BiSodCaseX.arc"))
    assert c.env_names == ["ZG", "ZD"]
    assert c.output_variables == ["CellMass", "Pressure", "Density", "Velocity", "NodeMass", "InternalEnergy",
                                  "PseudoViscosity", "Materiau"]
    assert c.time_history


@pytest.mark.parametrize("name,what", [("CAS_Compression_isotherme_ALU_monodimX", "MieGruneisen"),
                                       ("CAS_2D_ALE_Sedov_euler", "RemapALE"),
                                       ("CAS_2D_choc_plaque_crenau", "Tabulee")])
def test_unsupported_datasets_are_refused(name, what):
    with pytest.raises(ArcUnsupported, match=what):
        read_arc(os.path.join(DATASETS, name + ".arc"))


MINIMAL = """<?xml version='1.0'?>
<case codeversion="1.0" codename="Mahyco" xml:lang="en">
  <arcane><title>t</title><timeloop>MahycoLoop</timeloop></arcane>
  <mesh><meshgenerator><cartesian><nsd>1 1</nsd><origine>0.0 0.0</origine>
    <lx nx='8' prx='1.0'>1.</lx><ly ny='4' pry='1.0'>.5</ly></cartesian></meshgenerator></mesh>
  <mahyco>
    <material><name>A_mat</name></material>
    <environment><name>A</name><material>A_mat</material>
      <eos-model name="StiffenedGas"></eos-model></environment>
    <cas-model name="SOD"><cas-test>3</cas-test></cas-model>
    %s
  </mahyco>
</case>"""
TIMES = "<deltat-init>1e-5</deltat-init><deltat-min>1e-8</deltat-min><deltat-max>1e-2</deltat-max>" \
        "<final-time>.2</final-time>"


def test_defaults_of_the_axl_files():
    c = read_arc(MINIMAL % TIMES)
    assert (c.dim, c.n, c.length) == (2, (8, 4, 1), (1., .5, 0.))
    e = c.envs[0]
    assert (e.densite_initiale, e.pression_initiale, e.adiabatic_cst, e.limit_tension) == (0., 1e5, 1.4, .01)
    assert (e.linear_pseudo_coeff, e.quadratic_pseudo_coeff) == (.5, 1.2)
    assert (c.cfl, c.threshold, c.remap_threshold, c.ordre_projection, c.limiter) == (.3, 1e-10, 1e-10, 2, 0)
    assert c.with_projection and c.calcul_vitson_max and not c.conservation_energie_totale
    assert c.bcs == [] and c.longueur_caracteristique == 0


@pytest.mark.parametrize("extra,exc,match", [
    ("", ArcError, "deltat-init"),
    (TIMES + "<gravity>0. -9.81 0.</gravity>", ArcUnsupported, "gravity"),
    (TIMES + "<pseudo-centree>true</pseudo-centree>", ArcUnsupported, "pseudo-centree"),
    (TIMES + "<boundary-condition><surface>XMIN</surface><type>Pressure</type><value>1.</value>"
             "</boundary-condition>", ArcUnsupported, "Pressure"),
    (TIMES + "<remap name='RemapADI'><is-euler-scheme>Y</is-euler-scheme></remap>", ArcUnsupported, "euler"),
    (TIMES + "<cfl>abc</cfl>", ArcError, "cfl"),
])
def test_malformed_and_unsupported_options(extra, exc, match):
    with pytest.raises(exc, match=match):
        read_arc(MINIMAL % extra)
