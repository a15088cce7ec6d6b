"""Restart files in the reference's format (src/fv_restart.F90:474-613).

Fortran unformatted sequential records, each framed by its byte length as a
4-byte integer before and after the payload (the framing of the flang and
gfortran runtimes); integers are 4-byte, reals 8-byte, arrays in Fortran
(column-major) order.  The records are those write_restart_separate writes
for type_task = 3 and Vert_turb = 3 (d3_end):

    n
    time_jd, time_jd0, dt, dt_2D
    eta_n, eta_n_1, eta_n_2, U_n_2D, U_n_1, U_n_2
    U_n, V_n, U_rhsAB, V_rhsAB, Wvel
    bt, snu, Kv
    TF, SF, T_old, S_old
    Tclim, Sclim

``write`` takes the arrays from a model (fcg_get_state), ``hot_start`` puts
them back and completes the set-up as the reference does after
read_restart_separate (fcg_hot_start).
"""
from __future__ import annotations

import struct

import numpy as np

from . import gpu

RECORDS = (("eta_n", "eta_n_1", "eta_n_2", "U_n_2D", "U_n_1", "U_n_2"),
           ("U_n", "V_n", "U_rhsAB", "V_rhsAB", "Wvel"),
           ("bt", "snu", "Kv"),
           ("TF", "SF", "T_old", "S_old"),
           ("Tclim", "Sclim"))


def sizes(N, E, K):
    """Values per array for nod2D = N, elem2D = E, nsigma = K."""
    L = K - 1
    return {"eta_n": N, "eta_n_1": N, "eta_n_2": N, "U_n_2D": 2 * E, "U_n_1": 2 * E, "U_n_2": 2 * E,
            "U_n": L * E, "V_n": L * E, "U_rhsAB": L * E, "V_rhsAB": L * E, "Wvel": K * N,
            "bt": K * N, "snu": K * N, "Kv": K * N,
            "TF": L * N, "SF": L * N, "T_old": L * N, "S_old": L * N, "Tclim": L * N, "Sclim": L * N}


def _record(f, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def read_records(path):
    """The payloads of every record of a Fortran unformatted sequential file."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    i = 0
    while i < len(data):
        (n,) = struct.unpack_from("<i", data, i)
        payload = data[i + 4:i + 4 + n]
        (m,) = struct.unpack_from("<i", data, i + 4 + n)
        if m != n or len(payload) != n:
            raise ValueError(f"{path}: broken record framing at byte {i}")
        out.append(payload)
        i += 8 + n
    return out


def write_arrays(path, n, clock, arrays):
    """clock: (time_jd, time_jd0, dt, dt_2D); arrays: name -> FP64 array (Fortran order)."""
    with open(path, "wb") as f:
        _record(f, struct.pack("<i", n))
        _record(f, struct.pack("<4d", *clock))
        for names in RECORDS:
            _record(f, b"".join(np.ascontiguousarray(arrays[k], dtype="<f8").tobytes() for k in names))


def read(path, N, E, K):
    """-> (n, (time_jd, time_jd0, dt, dt_2D), {name: array})"""
    recs = read_records(path)
    if len(recs) != 2 + len(RECORDS):
        raise ValueError(f"{path}: {len(recs)} records, expected {2 + len(RECORDS)}")
    (n,) = struct.unpack("<i", recs[0])
    clock = struct.unpack("<4d", recs[1])
    size = sizes(N, E, K)
    arrays = {}
    for names, payload in zip(RECORDS, recs[2:]):
        v = np.frombuffer(payload, dtype="<f8")
        if v.size != sum(size[k] for k in names):
            raise ValueError(f"{path}: record {names} has {v.size} values")
        o = 0
        for k in names:
            arrays[k] = v[o:o + size[k]].copy()
            o += size[k]
    return n, clock, arrays


def write(path, model, n):
    """Restart of `model` after its internal step n (write_restart_separate)."""
    d = model.diagnostics()
    p = model.par
    arrays = {k: model.get(k) for names in RECORDS for k in names}
    write_arrays(path, n, (d["time_jd"], p.time_jd0, p.dt, p.dt_2D), arrays)


def hot_start(model, path):
    """read_restart_separate + the set-up that follows it; returns the
    restart's step count n (the reference's ini_time)."""
    n, clock, arrays = read(path, model.N, model.E, model.K)
    for k, v in arrays.items():
        model.set(k, v)
    model._ck(gpu.lib().fcg_hot_start(model.ctx, clock[0]), "fcg_hot_start")
    return n
