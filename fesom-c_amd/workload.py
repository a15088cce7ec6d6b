"""Named synthetic workloads as ready-to-step GPU models (Python host mirror).

``build_model("c2")`` is the 1M-node channel of the benchmark; ``meshgen``
holds the generators.  Extras (stratified T/S, uniform wind stress through the
reference's Large-Pond formula) are applied after ``initial_state`` as the
tests do.
"""
from __future__ import annotations

import numpy as np

from . import gpu, meshgen
from .runconfig import RunConfig


def build_model(name, scale=None, precision=0, closure=0):
    txt, mesh, ex = meshgen.config(name, scale=scale)
    cfg = RunConfig(txt)
    model = gpu.Model(cfg, mesh, precision=precision, closure=closure)
    if ex.get("T") is not None:
        model.set_tracers(ex["T"], ex["S"])
    if ex.get("wind") is not None:
        wx, wy = ex["wind"]
        tx, ty = meshgen.wind_stress(wx, wy)
        model.set("windx", np.full(mesh.nod2D, wx))
        model.set("windy", np.full(mesh.nod2D, wy))
        model.set("taux", np.full(mesh.elem2D, tx))
        model.set("tauy", np.full(mesh.elem2D, ty))
    return txt, mesh, ex, model


