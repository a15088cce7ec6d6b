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




def apply_member_forcing(model, meshes, extras, nodes=None, elems=None):
    """Uniform wind (Large-Pond stress) and initial T/S of each member, on the
    joined arrays of an ensemble context."""
    nodes = nodes or [m.nod2D for m in meshes]
    elems = elems or [m.elem2D for m in meshes]
    tau = [meshgen.wind_stress(*e["wind"]) for e in extras]
    model.set("windx", np.concatenate([np.full(n, e["wind"][0]) for n, e in zip(nodes, extras)]))
    model.set("windy", np.concatenate([np.full(n, e["wind"][1]) for n, e in zip(nodes, extras)]))
    model.set("taux", np.concatenate([np.full(n, t[0]) for n, t in zip(elems, tau)]))
    model.set("tauy", np.concatenate([np.full(n, t[1]) for n, t in zip(elems, tau)]))
    model.set_tracers(np.concatenate([e["T"] for e in extras]), np.concatenate([e["S"] for e in extras]))


class EnsembleSet:
    """An ensemble whose members have their own number of sigma levels: one
    fcg_create_ensemble context per level count (every stage one launch per
    class over all its members), stepped together.  ``classes`` come from
    meshgen.ensemble_classes; ``banks`` (from gpu.build_meshes, one per
    class) let several sets share the preprocessed meshes."""

    def __init__(self, classes, precision=0, closure=0, banks=None, net=None, serial=False):
        """``serial``: finish each class's step before the next class starts
        (no overlap between the classes' streams: per-kernel times without
        other kernels beside them)."""
        self.classes = classes
        self.serial = serial
        self.parts = []
        for i, (ns, txt, meshes, extras, ids) in enumerate(classes):
            cfg = RunConfig(txt)
            e = gpu.Ensemble(cfg, meshes, precision=precision, closure=closure,
                             prebuilt=None if banks is None else banks[i])
            apply_member_forcing(e, meshes, extras)
            if net is not None:
                e.attach_mlp(net)
            self.parts.append(e)
        self.N = sum(e.N for e in self.parts)
        self.E = sum(e.E for e in self.parts)
        self.D = sum(e.D for e in self.parts)
        self.members = sum(e.members for e in self.parts)
        self.node_levels = sum(e.N * e.L for e in self.parts)
        self.par = self.parts[0].par

    def step(self, n=1):
        for _ in range(int(n)):   # one step of every class in turn keeps every stream fed
            for e in self.parts:
                e.step(1)
                if self.serial:
                    e.sync()

    def sync(self):
        for e in self.parts:
            e.sync()

    def get(self, name):
        """The field of every class, concatenated in class order."""
        return np.concatenate([e.get(name) for e in self.parts])

    def diagnostics(self):
        ds = [e.diagnostics() for e in self.parts]
        return {"status": next((d["status"] for d in ds if d["status"]), 0),
                "energy": sum(d["energy"] for d in ds),
                "mlp_fallback": sum(d["mlp_fallback"] for d in ds)}

    def probe(self, kernel):
        for e in self.parts:
            e.probe(kernel)

    def probe_read(self):
        tot, n = 0.0, 0
        for e in self.parts:
            t, k = e.probe_read()
            tot += t
            n += k
        return tot, n

    def close(self):
        for e in self.parts:
            e.close()
        self.parts = []
