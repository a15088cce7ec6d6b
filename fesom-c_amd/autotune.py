"""Launch-shape autotuner.

For each tunable device kernel the candidate block sizes (64, 128, 256
threads; for the layer-parallel kernels 64 entities x 1, 2 or 4 layers per
block; 256, 512 or 1024 threads of whole columns for the wide-block kernels
-- the momentum right-hand side and the node-gather kernels of the tracer
advection, the vertical velocity and the node averaging, whose larger blocks
hold more node or element columns and so take more neighbours from LDS, at
the price of fewer resident blocks; and, for the kernels with a level-vector
variant, 1 or 4 (FP32) / 2 (FP64) layers per lane) are timed on the model's own mesh and
state with the library's HIP-event probe (fcg_probe), over a few internal
steps each, and the fastest is kept (fcg_set_launch).  The search advances
the model: run it during the warm-up of a run, or on a throw-away model of
the same mesh.  The resulting table can be stored and re-applied with
``apply``.  The tables of the benchmark workloads (one per level count of
the ensemble, both precisions) are written beside this module by

    python -c "import __graft_entry__ as g; g.load(); from fesomc_amd import autotune; autotune.main()" \
        --workload c5
"""
from __future__ import annotations

import json
import os

LAYER_KERNELS = ["k_tr_upwind", "k_tr_gradn_miura", "k_tr_miura", "k_tr_ab"]
COLUMN_KERNELS = ["k_jac_node"]
WIDE_KERNELS = ["k_velrhs", "k_tr_muscl", "k_tr_gradn_muscl", "k_vertvel_div", "k_el2nodes"]
# kernels with a level-vector variant: option "<kernel>:levels" = 1 or the
# vector width (4 layers per lane in FP32, 2 in FP64)
VECTOR_KERNELS = ["k_velrhs", "k_tr_muscl", "k_tr_gradn_muscl", "k_vertvel_div", "k_el2nodes", "k_velgrad3d",
                  "k_je", "k_update3d"]
# the tunable kernels the benchmark workloads launch (MUSCL tracers)
BENCH_KERNELS = ["k_tr_ab", "k_jac_node"] + WIDE_KERNELS + ["k_velgrad3d", "k_je", "k_update3d"]
CANDIDATES = (64, 128, 256)
WIDE_CANDIDATES = (256, 512, 1024)


def _cands(k, candidates):
    return WIDE_CANDIDATES if k in WIDE_KERNELS else candidates


def tune(model, kernels=None, candidates=CANDIDATES, steps=2, log=None):
    """Returns {kernel: (block, avg_ms)} and leaves the best shapes set.
    Advances the model by exactly tune_steps(kernels, candidates, steps)
    internal steps, whatever the timings (a wide-block candidate whose LDS
    does not fit is stepped with the default shape)."""
    from .gpu import GpuError
    table = {}
    for k in kernels or (LAYER_KERNELS + COLUMN_KERNELS + WIDE_KERNELS):
        best = None
        for bs in _cands(k, candidates):
            try:
                model.set_launch(k, bs)
            except GpuError:
                model.set_launch(k, 256)
                model.step(1 + steps)
                continue
            model.step(1)                      # first launches of the new shape, untimed
            model.probe(k)
            model.step(steps)
            model.probe(None)
            ms, n = model.probe_read()
            if n == 0:                         # not launched in this configuration
                continue
            t = ms / n
            if best is None or t < best[1]:
                best = (bs, t)
        if best is None:
            model.set_launch(k, 256)
            continue
        model.set_launch(k, best[0])
        table[k] = best
        if log:
            log(f"{k}: block {best[0]} ({best[1]:.3f} ms)")
    return table


def tune_steps(kernels=None, candidates=CANDIDATES, steps=2):
    return sum(len(_cands(k, candidates)) for k in kernels or (LAYER_KERNELS + COLUMN_KERNELS + WIDE_KERNELS)) \
        * (1 + steps)


def apply(model, table):
    for k, v in table.items():
        model.set_launch(k, v[0] if isinstance(v, (list, tuple)) else int(v))


def save(table, path):
    with open(path, "w") as f:
        json.dump({k: list(v) for k, v in table.items()}, f, indent=1)


def load(path):
    with open(path) as f:
        return {k: tuple(v) for k, v in json.load(f).items()}


def main():
    import argparse
    ap = argparse.ArgumentParser(description="tune the launch shapes of the internal step on one workload")
    ap.add_argument("--config", default="c2")
    ap.add_argument("--precision", type=int, default=1, help="0 FP64, 1 mixed")
    ap.add_argument("--closure", type=int, default=0, help="0 d3_end, 1 learned (needs --weights)")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    from .workload import build_model
    _, _, _, model = build_model(args.config, precision=args.precision, closure=args.closure)
    if args.weights:
        from . import closure
        model.attach_mlp(closure.load(args.weights))
    model.step(args.warmup)
    table = tune(model, steps=args.steps, log=lambda m: print(m, flush=True))
    model.close()
    save(table, args.out)


if __name__ == "__main__":
    main()


def tune_parts(model, log=None, steps=2):
    """Tune every part of an ensemble set (one table per level count) or a
    single model: {"nsigma": table}, each kernel timed alone on its part."""
    parts = model.parts if hasattr(model, "parts") else [model]
    out = {}
    for p in parts:
        out[str(p.K)] = tune(p, steps=steps, log=log)
    return out


def tune_set(model, kernels=None, candidates=CANDIDATES, steps=3, log=None, margin=0.01):
    """Tune by whole-step wall time: for an ensemble set whose level classes
    step on their own streams at the same time, a kernel's best shape alone
    is not its best shape beside the other classes' kernels (a wider block
    takes LDS the other streams' blocks would use).  For each kernel in turn
    every candidate is set on all parts and `steps` steps of the whole set
    are timed; the fastest is kept.  Returns {"nsigma": table} (the same
    shape for every class) and leaves it set."""
    import time
    parts = model.parts if hasattr(model, "parts") else [model]
    from .gpu import GpuError
    chosen = {}

    def timed():
        model.step(1)  # first launches of the shape (and the re-captured barotropic graph), untimed
        model.sync()
        t0 = time.perf_counter()
        model.step(steps)
        model.sync()
        return (time.perf_counter() - t0) / steps * 1e3

    def settings(k):
        """Candidate settings of kernel k, the current one first: block sizes
        of the one-level variant and, for a kernel with a level-vector
        variant, the vector variant (which runs whole columns in 256-thread
        blocks, so its block option does not apply)."""
        blocks = list(_cands(k, candidates)) if k in LAYER_KERNELS + COLUMN_KERNELS + WIDE_KERNELS else [None]
        out = []
        if k in VECTOR_KERNELS:
            levels = [parts[0].get_launch(k + ":levels")]
            levels += [v for v in (1, 2, 4) if v != levels[0]]
            keep = {k: parts[0].get_launch(k)} if blocks[0] else {}   # the one-level variant's shape, kept
            for lv in levels:
                if lv > 1:
                    out.append({k + ":levels": lv, **keep})
                else:
                    out += [{k + ":levels": 1, **({k: bs} if bs else {})} for bs in blocks]
        else:
            cur = parts[0].get_launch(k)
            out = [{k: bs} for bs in [cur] + [c for c in blocks if c != cur]]
        return out

    for k in kernels or (LAYER_KERNELS + COLUMN_KERNELS + WIDE_KERNELS):
        best = None
        # the current setting first; another replaces it only when it is at
        # least `margin` faster (whole-step times carry a little noise)
        for cand in settings(k):
            try:
                for p in parts:
                    for name, v in cand.items():
                        p.set_launch(name, v)
            except GpuError:          # a width that does not divide nsigma, or a block whose LDS does not fit
                continue
            t = timed()
            if best is None or t < best[1] * (1.0 - margin):
                best = (cand, t)
        if best is None:
            continue
        for p in parts:
            for name, v in best[0].items():
                p.set_launch(name, v)
        for name, v in best[0].items():
            chosen[name] = (v, best[1])
        if log:
            log(f"{k}: {best[0]} ({best[1]:.2f} ms per step)")
    return {str(p.K): dict(chosen) for p in parts}


def apply_parts(model, tables):
    """Apply a {"nsigma": table} set to the parts with that level count."""
    for p in (model.parts if hasattr(model, "parts") else [model]):
        t = tables.get(str(p.K))
        if t:
            apply(p, t)


def save_parts(tables, path):
    with open(path, "w") as f:
        json.dump({k: {kk: list(vv) for kk, vv in t.items()} for k, t in tables.items()}, f, indent=1)


def load_parts(path):
    with open(path) as f:
        return {k: {kk: tuple(vv) for kk, vv in t.items()} for k, t in json.load(f).items()}


def table_path(workload, precision):
    return os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        f"launch_{'mixed' if precision else 'fp64'}_{workload}.json")


def main():
    """Tune the benchmark workload (c5: the level-class ensemble; c2: the 1M-node
    channel) in both precisions and write the tables beside this module."""
    import argparse
    import sys
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--workload", choices=["c5", "c2"], default="c5")
    ap.add_argument("--steps", type=int, default=4)
    args = ap.parse_args()
    from . import closure, gpu, meshgen
    from .runconfig import RunConfig
    from .workload import EnsembleSet, build_model
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.workload == "c5":
        classes = meshgen.ensemble_classes("c4")
        banks = [gpu.build_meshes(RunConfig(c[1]), c[2]) for c in classes]
    net = closure.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "closure_c2.npz"))
    for prec, clos in ((1, gpu.FCG_CLOSURE_MLP), (0, gpu.FCG_CLOSURE_D3END)):
        if args.workload == "c5":
            model = EnsembleSet(classes, precision=prec, closure=clos, banks=banks,
                                net=net if clos == gpu.FCG_CLOSURE_MLP else None)
        else:
            model = build_model("c2", precision=prec, closure=clos)[3]
            if clos == gpu.FCG_CLOSURE_MLP:
                model.attach_mlp(net)
        model.step(2)
        tables = tune_set(model, kernels=BENCH_KERNELS, log=log, steps=args.steps)
        save_parts(tables, table_path(args.workload, prec))
        log(f"precision {prec}: {tables}")
        model.close()
    if args.workload == "c5":
        for b in banks:
            gpu.free_meshes(b)
    net.close()


if __name__ == "__main__":
    main()
