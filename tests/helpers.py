"""Shared helpers of the test suite: field lists and comparisons."""
import hashlib

import numpy as np

from mahyco_amd import types as T

CELL_FIELDS = [T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_SOUND_SPEED,
               T.MH_F_PSEUDO_VISCOSITY, T.MH_F_CELL_MASS, T.MH_F_FRACVOL, T.MH_F_CARACTERISTIC_LENGTH]
ENV_FIELDS = [T.MH_F_DENSITY, T.MH_F_PRESSURE, T.MH_F_INTERNAL_ENERGY, T.MH_F_CELL_MASS,
              T.MH_F_FRACVOL, T.MH_F_MASS_FRACTION]
NAMES = {v: k for k, v in T.CELL_FIELDS.items()}
NAMES.update({T.MH_N_VELOCITY: "Velocity", T.MH_N_NODE_MASS: "NodeMass"})


def fields(sim, nenv, dim):
    """All parity fields of a GpuCase / OracleCase as a dict name -> array."""
    out = {}
    for f in CELL_FIELDS:
        out[NAMES[f]] = sim.get(f)
    for k in range(nenv):
        for f in ENV_FIELDS:
            out[f"{NAMES[f]}[{k}]"] = sim.get(f, k)
    out["EnvMask"] = sim.get(T.MH_F_ENV_MASK)
    for c in range(dim):
        out[f"Velocity_{c}"] = sim.get(T.MH_N_VELOCITY, -1, c)
    out["NodeMass"] = sim.get(T.MH_N_NODE_MASS)
    return out


def max_rel(a, b):
    """max |a-b| / max(|b|, 1e-300), NaN-aware (NaN in both counts as equal)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.where(both_nan, 0.0, np.abs(a - b))
    s = np.maximum(np.abs(b), 1e-300)
    r = np.where(both_nan, 0.0, d / s)
    return float(np.nanmax(r)) if r.size else 0.0


def rel_err_scaled(a, b):
    """max |a-b| / max|b| over the field (for fields that cross zero, e.g. velocity)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b))) / scale


def bitwise_equal(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))


def digest(arr):
    """SHA-256 of a float64 array's bytes, every NaN written as the same quiet NaN (NaN
    payloads and signs differ between the host CPU and the GPU; the values are NaN on both)."""
    a = np.ascontiguousarray(arr, dtype=np.float64)
    a = np.where(np.isnan(a), np.float64(np.nan), a)
    return hashlib.sha256(a.tobytes()).hexdigest()


def summarize(sim, case):
    """Per parity field: digest, exact sum and the x-profile at j = k = 0 (hex floats)."""
    f = fields(sim, len(case.envs), case.dim)
    nx = case.cell_shape[-1]
    out = {}
    for k, v in f.items():
        row = v[:nx] if v.size == case.nb_cells else v[:nx + 1]
        out[k] = {"sha256": digest(v), "sum": float(np.sum(v)).hex(), "row": [float(x).hex() for x in row]}
    return out
