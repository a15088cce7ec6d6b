"""Benchmark: node-level-steps/s of the FESOM-C internal step on one MI355X.

Workload (the 1M-node channel): seeded 20x10 km channel, 1415x708 = 1,001,820
nodes, 90 % of quads split (~1.9M elements), 24 sigma levels (23 layers),
wetting/drying on (depth -2..30 m, 15 % intertidal), stratified T/S with
seeded smooth noise, seeded uniform wind, M2 tide at the open west boundary,
MUSCL tracers.  One node-level-step = one 2-D node x one layer advanced by one
internal step (including its Mt barotropic substeps).

Default (--path headline): the mixed-precision step with the learned closure
(weights fesom-c_amd/closure_c2.npz, trained on columns of the plain model by
fesomc_amd.closure) is timed, then the plain FP64 drop-in on the same
workload; the line reports both rates, the speed-up and the SSH/velocity/T/S
differences between the two after the same steps.  --path fp64 / mixed time
one path alone.

Each timed region holds exactly --steps internal steps with all inputs
resident in HBM, bracketed by synchronisation on both sides.  The probed
kernel is timed with HIP events on the library's stream (fcg_probe) and
reported against its algorithmic HBM bytes; traffic comes from the committed
copy-calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE passes.  The CPU baseline is
the FP64 oracle (scalar C, one core) on a bounded sample of the same physics.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def field_errors(model, ref, masks=None):
    """RMS and max differences of SSH, velocity and T/S between two runs, over
    all nodes/elements, or over the (node, element) masks given."""
    errs = {}
    for f, key, where in (("eta_n", "ssh_m", 0), ("U_n", "u_m_s", 1), ("V_n", "v_m_s", 1), ("TF", "t_degC", 0),
                          ("SF", "s_psu", 0)):
        a, b = model.get(f), ref.get(f)
        d = a - b
        if masks is not None:
            m = masks[where]
            d = d.reshape(m.size, -1)[m]
        errs[key + "_rms"] = float(np.sqrt(np.mean(d ** 2)))
        errs[key + "_max"] = float(np.max(np.abs(d)))
    return errs


def interior_masks(mesh, txt):
    """(node, element) masks of the part of the mesh outside the open-boundary
    sponge layer (ac == 1), elements in the reference's order (triangles
    first); None when the run has no sponge layer."""
    from fesomc_amd import gpu
    from fesomc_amd.runconfig import RunConfig
    ac = gpu.mesh_arrays(RunConfig(txt), mesh)["ac"]
    if np.all(ac == 1.0):
        return None
    nm = ac == 1.0
    tri = mesh.elem[:, 0] == mesh.elem[:, 3]
    order = np.concatenate([np.where(tri)[0], np.where(~tri)[0]])
    em = nm[mesh.elem[order] - 1].all(axis=1)
    return nm, em


def kernel_bytes(k, N, E, D, L, K, e3=8):
    """Algorithmic bytes of one launch: sum over distinct arrays read or
    written of entities x components x width (index arrays at 4 B; 3-D
    fields at e3 bytes: 8 in the FP64 path, 4 in the mixed-precision path;
    mesh geometry and 2-D fields stay FP64)."""
    i4 = 4
    LN, LE, KN, KE = L * N, L * E, K * N, K * E
    d8 = 8
    mesh_e = E * (4 * i4 + 4 * d8)          # elem nodes + w_cv
    edge_geo = i4 * (N + 1) + i4 * 2 * D + D * (8 + 8 + 32 + 16)   # n2e CSR, ednodes, edtri, ecross, edxdy
    if k == "k_tr_muscl":
        rd = edge_geo + d8 * E + N * (8 + 8 + 4) + e3 * (2 * LN + 4 * LN + 5 * LE + KN + 2 * LN + KN)
        return rd + e3 * 4 * LN      # reads: ab T/S, node gradients, U/V(filt), Je, W, T/S, Kv; writes T/S, T/S_old
    if k == "k_tr_gradn_muscl":
        return i4 * (N + 1) + i4 * 4 * E + E * (8 + 16 + 64) + e3 * 2 * LN + e3 * 4 * LN
    if k == "k_tr_upwind":
        rd = (i4 * (N + 1) + i4 * 2 * D + 2 * i4 * D * 2 + 4 * d8 * D + 3 * d8 * N + i4 * N
              + e3 * (2 * LN) + e3 * (3 * LE) + e3 * KN + e3 * (2 * LN) + e3 * KN)
        return rd + e3 * 4 * LN       # writes TF,SF,T_old,S_old
    if k == "k_velrhs":
        # element geometry (en, w_cv, gradient_sca, area, coriolis, elem_cos, mask, el_ied), node ac/atp/eta/a,
        # edge nodes/elements/cross/dxdy; U, V, AB history, Bpu/Bpv, 4 velocity gradients (E x L), W (N x K)
        rd = (mesh_e + E * 8 * d8 + 4 * d8 * E + 4 * i4 * E + d8 * N * 5 + e3 * (2 * LE + 2 * LE + 2 * LE + 4 * LE)
              + e3 * KN + i4 * D * 4 + d8 * D * 6)
        return rd + e3 * 4 * LE       # writes U_rhsAB, V_rhsAB, U_rhs, V_rhs
    if k == "k_d3end":
        rd = i4 * (N + 1) + i4 * 4 * E + d8 * E + e3 * 2 * LE + e3 * 2 * LN + e3 * 3 * KN + d8 * 5 * N
        return rd + e3 * 3 * KN       # writes bt, snu, Kv
    if k in ("k_vertvel", "k_vertvel_div"):
        return i4 * (N + 1) + i4 * 2 * D + i4 * 2 * D + d8 * 4 * D + d8 * N + e3 * 3 * LE + d8 * 2 * E + e3 * KN
    if k == "k_bpg":
        return mesh_e + d8 * E * 8 + d8 * 3 * N + d8 * KN + e3 * 2 * LE + d8 * 2 * E + d8 * E
    if k == "k_pressure":
        return d8 * N + e3 * 2 * LN + d8 * KN
    if k == "k_velgrad3d":
        return E * (4 * i4 * 3 + 8 * d8) + d8 * D * 2 + d8 * E + e3 * 2 * LE + e3 * 4 * LE
    if k == "k_update3d":
        return mesh_e + d8 * 3 * N + e3 * 2 * LE + e3 * LE + d8 * 4 * E + e3 * 4 * LE
    if k == "k_mlp_closure":
        # reads Unode, Vnode, TF, SF (L x N), a, eta, depth, ac (N); writes snu, Kv (K x N), the flag (N)
        return e3 * 4 * LN + d8 * 4 * N + e3 * 2 * KN + i4 * N
    raise KeyError(k)


MLP_FLOP_PER_SAMPLE = 2 * (5 * 128 + 128 * 128 + 128 * 128 + 128 * 2)   # 5 -> 128 -> 128 -> 128 -> 2


def pmc_traffic(kernel, path):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes over this workload (copy-calibrated), or None."""
    try:
        for line in open(path):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 5 and cells[0] == kernel:
                return float(cells[4])
    except OSError:
        pass
    return None


def cpu_baseline(budget_s=15.0):
    """FP64 oracle (scalar C, 1 core) on the same physics at 354x177 nodes."""
    import numpy as np
    import __graft_entry__ as ge
    ge.load()
    from fesomc_amd import meshgen
    from oracle.pyoracle import Oracle
    scale = (354, 177)
    txt, mesh, ex = meshgen.config("c2", scale=scale)
    o = Oracle(txt, mesh)
    o.set_tracers(ex["T"].reshape(-1), ex["S"].reshape(-1))
    wx, wy = ex["wind"]
    tx, ty = meshgen.wind_stress(wx, wy)
    o.set("windx", np.full(mesh.nod2D, wx))
    o.set("windy", np.full(mesh.nod2D, wy))
    o.set("taux", np.full(mesh.elem2D, tx))
    o.set("tauy", np.full(mesh.elem2D, ty))
    o.step(1)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 200:
            break
    dt = time.perf_counter() - t0
    N, L = o.size["N"], o.size["L"]
    return {"value": N * L * n / dt, "unit": "node-level-steps/s", "cores": 1, "kind": "port",
            "sample": f"FP64 oracle (scalar C restatement), config-2 physics on a {scale[0]}x{scale[1]} channel "
                      f"({N} nodes, {L} layers), {n} internal steps after 1 warm-up step, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--scale", default=None, metavar="NX,NY",
                    help="mesh nodes along and across the channel instead of the config's (same generator)")
    ap.add_argument("--probe", default="k_velrhs", help="kernel timed live for the roofline (the dominant one)")
    ap.add_argument("--error-steps", default="0", metavar="N|m2",
                    help="after the timed run, advance both the headline and the FP64 model (untimed) to N internal "
                         "steps in total ('m2': one M2 tidal period, 44712 s) and report their differences then")
    ap.add_argument("--path", choices=["headline", "mixed-mlp", "fp64", "mixed"], default="headline",
                    help="headline: mixed precision + learned closure, timed beside the plain FP64 drop-in")
    ap.add_argument("--closure-weights", default=os.path.join(ROOT, "fesom-c_amd", "closure_c2.npz"))
    ap.add_argument("--launch-table", default="auto",
                    help="block-size table from fesomc_amd.autotune (JSON); 'auto' = the stored table of this "
                         "config and precision if there is one, 'none' = library defaults")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (5 launches), the reference "
                         "kernel for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        x = torch.rand(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs one FESOM-C mesh partition on one GPU; --gpus {args.gpus} is not "
              "supported (the multi-GPU halo-exchange design is documented, not built)", file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__ as ge
    ge.load()
    from fesomc_amd import autotune, closure, gpu
    from fesomc_amd.workload import build_model

    held = {}

    def run(precision, clos, probe):
        t_setup = time.perf_counter()
        scale = tuple(int(v) for v in args.scale.split(",")) if args.scale else None
        txt, mesh, ex, model = build_model(args.config, scale=scale, precision=precision, closure=clos)
        held["mesh"], held["txt"] = mesh, txt
        if clos == gpu.FCG_CLOSURE_MLP:
            model.attach_mlp(closure.load(args.closure_weights))
        table = args.launch_table
        if table == "auto":
            table = os.path.join(ROOT, "fesom-c_amd", f"launch_{'mixed' if precision else 'fp64'}_{args.config}.json")
            table = table if os.path.exists(table) else "none"
        if table != "none":
            autotune.apply(model, autotune.load(table))
        setup = time.perf_counter() - t_setup
        model.step(args.warmup)
        model.sync()
        torch.cuda.synchronize()
        model.probe(probe)
        t0 = time.perf_counter()
        model.step(args.steps)
        model.sync()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        model.probe(None)
        k_ms, k_n = model.probe_read()
        diag = model.diagnostics()
        if diag["status"] != 0 or not math.isfinite(diag["energy"]):
            print(json.dumps({"error": "model state invalid", "diag": diag}), file=sys.stderr)
            sys.exit(1)
        return model, t1 - t0, k_ms / max(k_n, 1), k_n, setup, diag

    headline = args.path in ("headline", "mixed")   # paths compared against the FP64 drop-in
    prec = 0 if args.path == "fp64" else 1
    clos = gpu.FCG_CLOSURE_MLP if args.path in ("headline", "mixed-mlp") else gpu.FCG_CLOSURE_D3END
    model, sec, avg_ms, k_n, setup_s, diag = run(prec, clos, args.probe)
    N, E, D, L, K = model.N, model.E, model.D, model.L, model.K
    nls = N * L * args.steps / sec
    e3 = 4 if prec else 8
    kb = kernel_bytes(args.probe, N, E, D, L, K, e3)
    achieved = kb / (avg_ms * 1e-3) / 1e9
    tfile = os.path.join(ROOT, "profiles", f"pmc_traffic_{'mixed' if prec else 'fp64'}_step_c2.md")
    desc = {"headline": "mixed precision (FP32 3-D fluxes/tendencies; FP64 elevation/volume, wet/dry, BPG) "
                        "+ learned closure (BF16 MFMA MLP, d3_end fallback)",
            "fp64": "plain FP64 drop-in", "mixed": "mixed precision, d3_end closure",
            "mixed-mlp": "mixed precision + learned closure (not compared)"}[args.path]
    out = {
        "metric": "node-level-steps/s, mixed+learned vs plain FP64 GPU at <=1cm SSH RMS",
        "value": nls,
        "unit": "node-level-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sec / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if not prec else "f32/f64 (3-D fluxes f32, 2-D and BPG f64), mlp bf16",
        "data": "synthetic (seeded channel mesh generator, analytic forcing)",
        "config": {"workload": f"{args.config}: channel {N} nodes / {E} elements / {D} edges, nsigma {K}, "
                               f"Mt {model.par.Mt}, wet/dry, MUSCL tracers; {desc}",
                   "nodes": N, "elements": E, "layers": L, "Mt": model.par.Mt, "path": args.path},
        "roofline": {"kernel": args.probe, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": pmc_traffic(args.probe, tfile) if (args.config == "c2" and not args.scale) else None,
                     "avg_ms": avg_ms, "launches": k_n, "algorithmic_bytes": kb},
        "setup_s": setup_s,
        "energy": diag["energy"],
        "mlp_fallback_column_steps": diag["mlp_fallback"],
    }
    if headline:
        # the plain FP64 drop-in on the same workload: throughput and the error
        # of the headline path against it after the same warmup + timed steps
        ref, sec64, _, _, _, _ = run(0, gpu.FCG_CLOSURE_D3END, "")
        out["fp64_drop_in"] = {"value": N * L * args.steps / sec64, "ms_per_step": 1e3 * sec64 / args.steps}
        out["speedup_vs_fp64"] = sec64 / sec
        interior = interior_masks(held["mesh"], held["txt"])
        out["error_vs_fp64_after_steps"] = field_errors(model, ref)
        total = args.warmup + args.steps
        target = (int(math.ceil(44712.0 / model.par.dt)) if args.error_steps == "m2" else int(args.error_steps))
        if target > total:
            # both runs advance in lock-step chunks, each chunk reported on stderr
            t_err = time.perf_counter()
            while total < target:
                n = min(500, target - total)
                model.step(n)
                ref.step(n)
                model.sync()
                ref.sync()
                total += n
                print(f"error leg: {total}/{target} steps, {time.perf_counter() - t_err:.0f} s", file=sys.stderr,
                      flush=True)
            d_m, d_r = model.diagnostics(), ref.diagnostics()
            out["error_vs_fp64_long"] = {"steps": total, "model_time_s": total * model.par.dt,
                                         **field_errors(model, ref),
                                         "outside_sponge": field_errors(model, ref, interior) if interior else None,
                                         "status": [d_m["status"], d_r["status"]],
                                         "mlp_fallback_column_steps": d_m["mlp_fallback"]}
        ref.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    model.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
