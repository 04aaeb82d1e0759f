"""Restart files: the solver state written and read in flowPsi's file set.

Files in a restart directory (src/solverRestart.loci:116-164,
src/solverScalarTransport.loci:303-341):
  flowVars_<case>       ncycle, stime, Pambient, temperature, gagePressure,
                        u, cellcenter
  flowVarsPrev_<case>   time level n-1 (BDF2): temperature, gagePressure, u,
                        Pambient
  <X>_<case>            a turbulence scalar X (nuTilde; k and w)
  <X>Prev_<case>        its level n-1 (BDF2)
Reading follows src/solverIC.loci:572-660 and
src/solverScalarTransport.loci:226-296: the gage pressure is shifted by the
file's Pambient minus the current one, and a missing Prev file means
q{n-1} = q{n}.

Each variable is an HDF5 group holding a dataset "data" (float64 [nc], a
compound {x, y, z} [nc] for vectors, [1] for parameters) and an
"Interval Set" dataset with the domain [[0, nc - 1]], the container form
that the VOG files' cell_info groups use.  Loci's container writer is not
part of the reference, so reading files that Loci wrote is unverified.
Writing goes through libhdf5 (flowpsi_amd/h5lib.py); reading through
flowpsi_amd/hdf5.py.
"""
from __future__ import annotations

import os

import numpy as np

from . import h5lib, hdf5

SCALARS = {"none": [], "sa": ["nuTilde"], "sst": ["k", "w"]}


def _write_file(path, vars_):
    with h5lib.Writer(path) as w:
        for name, (kind, a) in vars_.items():
            w.group(name)
            n = 1 if kind == "param" else len(a)
            w.dataset(f"{name}/Interval Set", np.array([[0, n - 1]], np.int64))
            if kind == "vect3d":
                w.dataset(f"{name}/data", np.ascontiguousarray(a, np.float64), vect3d=True)
            else:
                w.dataset(f"{name}/data", np.atleast_1d(np.asarray(a, np.float64 if kind != "int" else np.int64)))


def _read_var(f, name):
    a = f[f"{name}/data"]
    if a.dtype.names:
        return np.stack([a["x"], a["y"], a["z"]], axis=1)
    return a


def write(dirpath, case, solver, ncycle=0, stime=0.0):
    """Write the solver's state (and, time-accurate, level n-1) as restart
    files of case `case` under dirpath."""
    prm = solver.prm
    qm1 = None
    if prm.get("unsteady", False) and prm.get("time_scheme", "bdf2") == "bdf2":
        _, qm1 = solver.get_time_levels()
    write_state(dirpath, case, solver.grid, prm, solver.get_state(), qm1, ncycle, stime)


def write_state(dirpath, case, g, prm, q, qm1=None, ncycle=0, stime=0.0):
    """Write a state q [nvar][nc] (and level n-1, qm1) of grid g as restart
    files of case `case` under dirpath."""
    os.makedirs(dirpath, exist_ok=True)
    pamb = float(prm.get("Pambient", 0.0))
    cc = np.ascontiguousarray(g.cc.T)
    _write_file(os.path.join(dirpath, f"flowVars_{case}"), {
        "ncycle": ("int", [ncycle]), "stime": ("param", [stime]), "Pambient": ("param", [pamb]),
        "temperature": ("store", q[0]), "gagePressure": ("store", q[4]), "u": ("vect3d", q[1:4].T),
        "cellcenter": ("vect3d", cc)})
    names = SCALARS[prm.get("turb", "none")]
    for j, x in enumerate(names):
        _write_file(os.path.join(dirpath, f"{x}_{case}"), {x: ("store", q[5 + j])})
    if qm1 is not None:
        _write_file(os.path.join(dirpath, f"flowVarsPrev_{case}"), {
            "temperature": ("store", qm1[0]), "gagePressure": ("store", qm1[4]), "u": ("vect3d", qm1[1:4].T),
            "Pambient": ("param", [pamb])})
        for j, x in enumerate(names):
            _write_file(os.path.join(dirpath, f"{x}Prev_{case}"), {x: ("store", qm1[5 + j])})


def _state_from(dirpath, stem, case, names, pamb, nc):
    f = hdf5.File(os.path.join(dirpath, f"{stem}_{case}"))
    T = _read_var(f, "temperature")
    pg = _read_var(f, "gagePressure") + (float(_read_var(f, "Pambient")[0]) - pamb)
    u = _read_var(f, "u")
    if len(T) != nc:
        raise ValueError(f"{stem}_{case} holds {len(T)} cells, the grid {nc}")
    rows = [T, u[:, 0], u[:, 1], u[:, 2], pg]
    suffix = "Prev" if stem.endswith("Prev") else ""
    for x in names:
        fx = hdf5.File(os.path.join(dirpath, f"{x}{suffix}_{case}"))
        rows.append(_read_var(fx, x))
    return np.vstack(rows), f


def read(dirpath, case, solver):
    """Set the solver's state (and level n-1) from restart files; returns
    {ncycle, stime}."""
    prm = solver.prm
    names = SCALARS[prm.get("turb", "none")]
    pamb = float(prm.get("Pambient", 0.0))
    nc = solver.grid.nc
    q, f = _state_from(dirpath, "flowVars", case, names, pamb, nc)
    solver.set_state(q)
    prev = os.path.join(dirpath, f"flowVarsPrev_{case}")
    if prm.get("unsteady", False) and os.path.exists(prev):
        qm1, _ = _state_from(dirpath, "flowVarsPrev", case, names, pamb, nc)
        solver.set_time_levels(None, qm1)
    return {"ncycle": int(_read_var(f, "ncycle")[0]), "stime": float(_read_var(f, "stime")[0])}
