"""Readers and writers for FESOM-C's mesh and forcing files.

Formats follow the reference's serial reader ``read_mesh_ser``
(src/fv_main.F90:787-945) and the tidal boundary reader in
``initial_state`` (src/fv_ini.F90:204-240):

* ``nod2d.out``  first line ``nod2D``, then ``n x y index`` per node.
* ``elem2d.out`` first line ``elem2D``, then 4 node ids per element (a
  triangle repeats its first node); a 3-column file means triangles only.
* ``depth.out``  ``nod2D`` values, positive down.
* ``m2.out``     number of open-boundary nodes, then ``idx amp phase ...``.

Files may be gzip-compressed (``.gz`` suffix).
"""
from __future__ import annotations

import gzip
import os
from dataclasses import dataclass, field

import numpy as np


def _open(path):
    if path.endswith(".gz"):
        return gzip.open(path, "rt")
    if not os.path.exists(path) and os.path.exists(path + ".gz"):
        return gzip.open(path + ".gz", "rt")
    return open(path, "r")


@dataclass
class RawMesh:
    """Mesh exactly as the reference files hold it (1-based element node ids)."""

    xy: np.ndarray          # (N, 2) float64, file units (metres or degrees)
    index: np.ndarray       # (N,) int32, 0 interior / 1 solid / 2 open boundary
    elem: np.ndarray        # (E, 4) int32, 1-based, triangle: col 3 == col 0
    depth: np.ndarray       # (N,) float64, positive down
    ob: np.ndarray | None = None   # (nobn, 2*n_harmonics) amp/phase rows of m2.out
    meta: dict = field(default_factory=dict)

    @property
    def nod2D(self) -> int:
        return int(self.xy.shape[0])

    @property
    def elem2D(self) -> int:
        return int(self.elem.shape[0])


def read_nod2d(path):
    with _open(path) as f:
        n = int(f.readline().split()[0])
        data = np.loadtxt(f, max_rows=n)
    data = np.atleast_2d(data)
    ids = data[:, 0].astype(np.int64)
    xy = np.zeros((n, 2))
    idx = np.zeros(n, dtype=np.int32)
    # nodes are stored by their id column, as the reference does (index_nod2D(nq)=ind)
    xy[ids - 1, 0] = data[:, 1]
    xy[ids - 1, 1] = data[:, 2]
    idx[ids - 1] = data[:, 3].astype(np.int32)
    return xy, idx


def read_elem2d(path):
    with _open(path) as f:
        e = int(f.readline().split()[0])
        toks = f.read().split()
    vals = np.array(toks, dtype=np.int64)
    if vals.size >= 4 * e:
        elem = vals[: 4 * e].reshape(e, 4)
    else:  # triangles only, 3 columns
        tri = vals[: 3 * e].reshape(e, 3)
        elem = np.concatenate([tri, tri[:, :1]], axis=1)
    return elem.astype(np.int32)


def read_depth(path, n):
    with _open(path) as f:
        vals = np.array(f.read().split(), dtype=np.float64)
    return vals[:n]


def read_m2(path, n_harm):
    with _open(path) as f:
        n = int(f.readline().split()[0])
        rows = []
        for _ in range(n):
            t = f.readline().split()
            rows.append([float(x) for x in t[1:1 + 2 * n_harm]])
    return np.array(rows, dtype=np.float64).reshape(n, 2 * n_harm)


def read_mesh(meshpath, n_harm=0, with_depth=True):
    xy, idx = read_nod2d(os.path.join(meshpath, "nod2d.out"))
    elem = read_elem2d(os.path.join(meshpath, "elem2d.out"))
    depth = read_depth(os.path.join(meshpath, "depth.out"), xy.shape[0]) if with_depth else np.zeros(xy.shape[0])
    ob = None
    m2 = os.path.join(meshpath, "m2.out")
    if n_harm and (os.path.exists(m2) or os.path.exists(m2 + ".gz")):
        ob = read_m2(m2, n_harm)
    return RawMesh(xy=xy, index=idx, elem=elem, depth=depth, ob=ob)


def write_mesh(meshpath, mesh: RawMesh, fmt_xy="%.10f"):
    os.makedirs(meshpath, exist_ok=True)
    n = mesh.nod2D
    with open(os.path.join(meshpath, "nod2d.out"), "w") as f:
        f.write(f"{n}\n")
        for i in range(n):
            f.write(("%d " + fmt_xy + " " + fmt_xy + " %d\n") % (i + 1, mesh.xy[i, 0], mesh.xy[i, 1], mesh.index[i]))
    with open(os.path.join(meshpath, "elem2d.out"), "w") as f:
        f.write(f"{mesh.elem2D}\n")
        np.savetxt(f, mesh.elem, fmt="%d")
    with open(os.path.join(meshpath, "depth.out"), "w") as f:
        np.savetxt(f, mesh.depth, fmt="%.10f")
    if mesh.ob is not None:
        obn = np.nonzero(mesh.index == 2)[0]
        with open(os.path.join(meshpath, "m2.out"), "w") as f:
            f.write(f"{len(obn)}\n")
            for k, node in enumerate(obn):
                f.write(f"{node + 1} " + " ".join("%.6f" % v for v in mesh.ob[k]) + "\n")
