"""Launch-shape autotuner.

For each tunable device kernel the candidate block sizes (64, 128, 256
threads; for the layer-parallel kernels 64 entities x 1, 2 or 4 layers per
block; 256, 512 or 1024 threads of whole columns for the wide-block kernels,
whose larger blocks take more neighbours from LDS) are timed on the model's own mesh and state with the library's
HIP-event probe (fcg_probe), over a few internal steps each, and the fastest
is kept (fcg_set_launch).  The search advances the model: run it during the
warm-up of a run, or on a throw-away model of the same mesh.  The resulting
table can be stored and re-applied with ``apply``.
"""
from __future__ import annotations

import json

LAYER_KERNELS = ["k_tr_upwind", "k_tr_gradn_miura", "k_tr_miura", "k_tr_ab"]
COLUMN_KERNELS = ["k_d3end", "k_jac_node"]
WIDE_KERNELS = ["k_velrhs"]
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
    ap.add_argument("--steps", type=int, default=2)
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
