"""Arcane dataset (.arc) reader and runner for the options of this path.

read_arc() turns a MaHyCo dataset into a cases.Case: the <arcane> time loop, the Cartesian
mesh generator, and the <mahyco> options of Mahyco.axl, RemapADI.axl, the PerfectGas,
StiffenedGas and Fictif EOS services and the SOD, SEDOV, POINTTRIPLE, RIDER and ADVECTION
case services, with the .axl defaults for what the dataset leaves out (Mahyco.axl:263-557,
RemapADI.axl:165-209, eos/*/*.axl, casTest/*/*.axl).  A dataset that asks for anything this
path does not run (another EOS, elasto-plasticity, another remap service or case service,
gravity, anti-drift, other boundary-condition types, a non-uniform mesh ...) raises
ArcUnsupported naming the option, rather than running something else.

    python -m mahyco_amd.arc Donnees.arc [--out DIR] [--max-iterations N]

runs a dataset on the GPU to its final time and writes what the reference's run writes:
the EnSight Gold post-processing of the <output> variables (output/depouillement) and,
with the TimeHistory module, the EnvSummation curves (output/courbes/gnuplot).
"""
import argparse
import os
import sys
import xml.etree.ElementTree as ET

from . import cases
from . import types as T


class ArcError(ValueError):
    """The dataset is malformed (missing required option, unreadable value)."""


class ArcUnsupported(ArcError):
    """The dataset asks for a feature this path does not run."""


CASE_MODELS = {"SOD": T.MH_CASE_SOD, "SEDOV": T.MH_CASE_SEDOV, "POINTTRIPLE": T.MH_CASE_POINTTRIPLE,
               "RIDER": T.MH_CASE_RIDER, "ADVECTION": T.MH_CASE_ADVECTION}
LENGTHS = {"racine-cubique-volume": 0, "faces-opposees": 1, "monodimX": 2, "monodimY": 3, "monodimZ": 4}
EOS = {"PerfectGas": T.MH_EOS_PERFECT_GAS, "StiffenedGas": T.MH_EOS_STIFFENED_GAS, "Fictif": T.MH_EOS_FICTIF}
# limit-tension defaults of the EOS services (PerfectGasEOS.axl, StiffenedGasEOS.axl, FictifEOS.axl)
LIMIT_TENSION_DEFAULT = {"PerfectGas": 0.0, "StiffenedGas": 0.01, "Fictif": 0.0}
# rayon-bulle defaults of the bubble case services (RIDER.axl, ADVECTION.axl)
RAYON_BULLE_DEFAULT = {"RIDER": 0.15, "ADVECTION": 0.25}


def _text(e, tag):
    c = e.find(tag)
    return None if c is None or c.text is None else c.text.strip()


def _real(e, tag, default=None):
    t = _text(e, tag)
    if t is None:
        if default is None:
            raise ArcError(f"<{tag}> is required")
        return float(default)
    try:
        return float(t)
    except ValueError:
        raise ArcError(f"<{tag}>: not a real: {t!r}") from None


def _int(e, tag, default):
    t = _text(e, tag)
    if t is None:
        return default
    try:
        return int(t)
    except ValueError:
        raise ArcError(f"<{tag}>: not an integer: {t!r}") from None


def _bool(e, tag, default):
    t = _text(e, tag)
    if t is None:
        return default
    if t.lower() in ("true", "1"):
        return True
    if t.lower() in ("false", "0"):
        return False
    raise ArcError(f"<{tag}>: not a boolean: {t!r}")


def _real3(e, tag, default=(0., 0., 0.)):
    t = _text(e, tag)
    if t is None:
        return tuple(default)
    v = [float(x) for x in t.split()]
    if not 1 <= len(v) <= 3:
        raise ArcError(f"<{tag}>: expected up to three reals: {t!r}")
    return tuple(v + [0.] * (3 - len(v)))


def _inert(e, tag, default, what):
    """An option whose non-default value selects something outside this path."""
    t = _text(e, tag)
    if t is None:
        return
    if isinstance(default, bool):
        same = _bool(e, tag, default) == default
    else:
        try:
            same = float(t) == float(default)
        except ValueError:
            same = t.lower() == str(default).lower()
    if not same:
        raise ArcUnsupported(f"<{tag}>{t}</{tag}>: {what}")


def _mesh(root):
    gen = root.find("mesh/meshgenerator")
    if gen is None or gen.find("cartesian") is None:
        raise ArcUnsupported("only the Cartesian mesh generator (<meshgenerator><cartesian>) is read")
    cart = gen.find("cartesian")
    origin = _real3(cart, "origine")
    n, length = [1, 1, 1], [0., 0., 0.]
    axes = 0
    for d, (tag, attr, pr) in enumerate((("lx", "nx", "prx"), ("ly", "ny", "pry"), ("lz", "nz", "prz"))):
        e = cart.find(tag)
        if e is None:
            if d < 2:
                raise ArcError(f"<{tag}> is required")
            continue
        axes += 1
        if float(e.get(pr, "1.0")) != 1.0:
            raise ArcUnsupported(f"<{tag} {pr}='{e.get(pr)}'>: only uniform spacing")
        n[d] = int(e.get(attr))
        length[d] = float(e.text.strip())
    return axes, tuple(n), origin, tuple(length)


def _environment(e):
    name = _text(e, "name") or "?"
    if len(e.findall("material")) != 1:
        raise ArcUnsupported(f"environment {name}: one material per environment")
    eos = e.find("eos-model")
    if eos is None:
        raise ArcError(f"environment {name}: <eos-model> is required")
    kind = eos.get("name")
    if kind not in EOS:
        raise ArcUnsupported(f"environment {name}: eos-model {kind} (PerfectGas, StiffenedGas and Fictif are run)")
    elasto = e.find("elasto-model")
    if elasto is not None and elasto.get("name") not in (None, "NoModel"):
        raise ArcUnsupported(f"environment {name}: elasto-model {elasto.get('name')}")
    if e.find("energy-depot") is not None:
        raise ArcUnsupported(f"environment {name}: energy-depot")
    _inert(e, "initialisation-utilisateur", False, "user initialisation")
    _inert(eos, "tension-damage-thresold", -1.e12, "damage (Endommagement)")
    _inert(eos, "density-damage-thresold", 0., "damage (Endommagement)")
    _inert(eos, "formulation", "VE", "EOS formulation")
    _inert(eos, "init-formulation", "VP", "EOS initial formulation")
    env = cases.Env(densite_initiale=_real(e, "densite-initiale", 0.),
                    pression_initiale=_real(e, "pression-initiale", 1.e5),
                    adiabatic_cst=_real(eos, "adiabatic-cst", 1.4), eos=EOS[kind],
                    limit_tension=_real(eos, "limit-tension", LIMIT_TENSION_DEFAULT[kind]),
                    specific_heat=_real(eos, "specific-heat", 0.),
                    linear_pseudo_coeff=_real(e, "linear-pseudo-coeff", .5),
                    quadratic_pseudo_coeff=_real(e, "quadratic-pseudo-coeff", 1.2),
                    vitesse_initiale=_real3(e, "vitesse-initiale"),
                    energie_initiale=_real(e, "energie-initiale", 0.))
    if kind == "Fictif":
        env.fictif_tdebut = _real(eos, "tdebut-pression", 0.)
        env.fictif_tfin = _real(eos, "tfin-pression", 0.)
        env.fictif_p0 = _real(eos, "valeur-debut-pression", 0.)
        env.fictif_dpdt = _real(eos, "valeur-dependance-temps", 0.)
    return name, env


def read_arc(source):
    """The cases.Case of an Arcane dataset (a path, or the XML text itself).  The case's
    env_names attribute holds the environment names, in dataset order."""
    if isinstance(source, str) and source.lstrip().startswith("<"):
        root = ET.fromstring(source)
    else:
        root = ET.parse(source).getroot()
    title = _text(root, "arcane/title") or "dataset"
    loop = _text(root, "arcane/timeloop")
    if loop not in ("MahycoLoop", "MahycoLagrangeLoop"):
        raise ArcUnsupported(f"time loop {loop!r} (MahycoLoop and MahycoLagrangeLoop are run)")
    axes, n, origin, length = _mesh(root)
    m = root.find("mahyco")
    if m is None:
        raise ArcError("<mahyco> is required")
    named = [_environment(e) for e in m.findall("environment")]
    if not named:
        raise ArcError("at least one <environment> is required")
    if len(named) > 4:
        raise ArcUnsupported(f"{len(named)} environments (at most 4)")
    cm = m.find("cas-model")
    model = "EXTERNAL" if cm is None else cm.get("name", "EXTERNAL")
    if model not in CASE_MODELS:
        raise ArcUnsupported(f"cas-model {model} (SOD, SEDOV, POINTTRIPLE, RIDER and ADVECTION are run)")
    remap = m.find("remap")
    rname = "RemapADI" if remap is None else remap.get("name", "RemapADI")
    if rname != "RemapADI":
        raise ArcUnsupported(f"remap service {rname} (RemapADI is run)")
    if remap is None:
        remap = ET.Element("remap")
    _inert(remap, "is-euler-scheme", "Full", "is-euler-scheme other than Full")
    _inert(remap, "calcul_flux_masse", 0, "calcul_flux_masse")
    _inert(m, "pseudo-centree", False, "pseudo-centree")
    _inert(m, "coef-antiderive", 0., "anti-drift correction")
    _inert(m, "deltat-constant", False, "constant time step")
    if _real3(m, "gravity") != (0., 0., 0.):
        raise ArcUnsupported("gravity")
    with_projection = _bool(m, "with-projection", True)
    if loop == "MahycoLagrangeLoop" and with_projection:
        raise ArcUnsupported("MahycoLagrangeLoop with with-projection true")
    lc = _text(m, "longueur-caracteristique") or "racine-cubique-volume"
    if lc not in LENGTHS:
        raise ArcUnsupported(f"longueur-caracteristique {lc}")
    bcs = []
    for b in m.findall("boundary-condition"):
        surf, typ = _text(b, "surface"), _text(b, "type")
        if surf not in T.SURFACES:
            raise ArcUnsupported(f"boundary-condition surface {surf}")
        if typ not in T.BC_TYPES:
            raise ArcUnsupported(f"boundary-condition type {typ} (Vx, Vy and Vz are run)")
        bcs.append((surf, typ, _real(b, "value")))
    if len(bcs) > 12:
        raise ArcUnsupported(f"{len(bcs)} boundary conditions (at most 12)")
    c = cases.Case(title, 3 if axes == 3 else 2, n, origin, length, [e for _, e in named], CASE_MODELS[model],
                   _int(cm, "cas-test", -1) if cm is not None else -1,
                   with_projection=loop == "MahycoLoop" and with_projection,
                   cfl=_real(m, "cfl", .3), threshold=_real(m, "threshold", 1.e-10),
                   deltat_init=_real(m, "deltat-init"), deltat_min=_real(m, "deltat-min"),
                   deltat_max=_real(m, "deltat-max"), final_time=_real(m, "final-time"),
                   calcul_vitson_max=_bool(m, "calcul-vitson-max", True), bcs=bcs)
    if axes == 2:
        c.length = (length[0], length[1], 0.)
    c.ordre_projection = _int(remap, "ordre-projection", 2)
    c.limiter = _int(remap, "projection-limiteur-id", 0)
    c.limiter_pure = _int(remap, "projection-limiteur-pure-id", 0)
    c.remap_threshold = _real(remap, "threshold", 1.e-10)
    c.pente_borne = _bool(remap, "projection-pente-borne", False)
    c.debar_fix = _int(remap, "projection-pente-borne-debar-fix", 0)
    c.pente_borne_mixte = _bool(remap, "projection-pente-borne-mixte", False)
    c.simple_pente = _bool(remap, "projection-simple-pente", False)
    c.conservation_energie_totale = _bool(remap, "conservation-energie-totale", False)
    c.with_newton = _bool(m, "with-newton", False)
    c.pression_explicite = _bool(m, "pression-explicite", False)
    c.schema_csts = _bool(m, "schema-csts", False)
    c.sans_lagrange = _bool(m, "sans-lagrange", False)
    c.longueur_caracteristique = LENGTHS[lc]
    if model in RAYON_BULLE_DEFAULT:
        c.rider_reverse = _bool(cm, "reverse-option", False)
        c.rider_parametre = _real(cm, "parametre", 0.)
        c.rider_rayon_bulle = _real(cm, "rayon-bulle", RAYON_BULLE_DEFAULT[model])
        c.rider_densite_lineaire = _bool(cm, "densite-interieure-initiale-lineaire", False)
    c.env_names = [name for name, _ in named]
    out = root.find("arcane-post-processing/output")
    c.output_variables = [v.text.strip() for v in out.findall("variable")] if out is not None else []
    th = root.find("arcane/modules")
    c.time_history = any(mod.get("name") == "TimeHistory" and mod.get("active", "true") == "true"
                         for mod in (th.findall("module") if th is not None else []))
    return c


def run_arc(path, out_dir="output", max_iterations=None, graph=6):
    """Runs the dataset on the GPU to its final time (or max_iterations) and writes the
    EnSight post-processing and the EnvSummation curves under out_dir; returns the case's
    final status, iteration and time."""
    from .lib import GpuCase
    case = read_arc(path)
    g = GpuCase(case)
    if case.time_history:
        g.history_enable(10, 1 << 16)
    st, left = T.MH_OK, max_iterations
    while st == T.MH_OK and (left is None or left > 0):
        step = 500 if left is None else min(500, left)
        st = g.run(step, graph=graph)
        if left is not None:
            left -= step
    if st < 0:
        g.close()
        raise RuntimeError(f"{case.name}: {T.STATUS_NAMES.get(st, st)} at iteration {g.iteration}")
    if case.output_variables:
        g.ensight_write(os.path.join(out_dir, "depouillement"), case.output_variables)
    if case.time_history:
        g.history_write_gnuplot(os.path.join(out_dir, "courbes", "gnuplot"), case.env_names)
    res = (st, g.iteration, g.time)
    g.close()
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("dataset")
    ap.add_argument("--out", default="output")
    ap.add_argument("--max-iterations", type=int, default=None)
    args = ap.parse_args(argv)
    try:
        st, it, t = run_arc(args.dataset, args.out, args.max_iterations)
    except ArcUnsupported as e:
        print(f"unsupported: {e}", file=sys.stderr)
        return 2
    print(f"iteration={it} time={t!r} status={T.STATUS_NAMES.get(st, st)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
