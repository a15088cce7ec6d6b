"""Benchmark: node-level-steps/s of the FESOM-C internal step on one MI355X.

Default workload (--workload c5): the benchmark's ensemble -- 64 seeded
channels (20x10 km) of 50k..400k nodes each, 14.93M nodes and 28.4M elements
in all, with their own number of sigma levels (16, 20, 24, 28 or 32; 330.7M
node-levels), tidal amplitude (+-10 %) and wind (+-3 m/s per component);
wetting/drying, MUSCL tracers, stratified T/S with seeded smooth noise, 3 km
open-boundary sponge, dt 5.5 s, 10 barotropic substeps.  Members that share a
level count share one context (one launch per stage over all of them); the
five contexts step together.  The headline path is mixed precision with the
learned closure in every member (weights fesom-c_amd/closure_c2.npz, trained
on columns of the plain 1M-node channel), timed against the plain FP64
drop-in on the same ensemble.  --workload c2 times the single 1M-node channel
instead (1415x708 nodes, 24 levels).  One node-level-step = one 2-D node x one
layer advanced by one internal step (including its barotropic substeps).

Each timed region holds exactly --steps internal steps with all inputs
resident in HBM, bracketed by synchronisation on both sides.  The dominant
kernel (--probe) and a second one (--probe2) are timed with HIP events on the
library's streams over two further steps each, with the ensemble's level
classes stepped in turn (in the timed region their streams overlap, and an
event pair on one stream would also bracket the other streams' kernels).  The line reports the learned closure against the
BF16 matrix-core peak and the momentum kernel against the HBM peak, the
errors of the headline path against the FP64 run after the same steps, and
the CPU baseline: the FP64 oracle (scalar C, one core) on the CPU
configuration (the 5.1k-node channel).
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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(budget_s=20.0):
    """The FP64 oracle (scalar C restatement, one core) on the CPU
    configuration: the 5.1k-node channel, 10 sigma levels, upwind tracers,
    M2 tide, timed over internal steps after one warm-up step."""
    import __graft_entry__ as ge
    ge.load()
    from fesomc_amd import meshgen
    from oracle.pyoracle import Oracle
    txt, mesh, ex = meshgen.config("c1")
    o = Oracle(txt, mesh)
    o.step(1)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 1000:
            break
    dt = time.perf_counter() - t0
    N, L = o.size["N"], o.size["L"]
    return {"value": N * L * n / dt, "unit": "node-level-steps/s", "cores": 1, "kind": "port",
            "host_cpu": cpu_model(), "host_nproc": os.cpu_count(),
            "sample": f"FP64 oracle (scalar C restatement, one thread) on the CPU configuration: c1 channel "
                      f"{N} nodes, {L} layers, upwind tracers, M2 tide; {n} internal steps after 1 warm-up step, "
                      f"{dt:.1f} s"}


def mlp_samples(model_parts):
    """Samples of the learned closure per step: every level of every column
    outside the open-boundary sponge layer (a lower bound: groups of columns
    that touch the layer run the network on all of theirs)."""
    tot = 0
    for part in model_parts:
        k = part.K
        for v in getattr(part, "views", None) or [part.view]:
            ac = np.ctypeslib.as_array(v.ac, shape=(v.nod2D,))
            tot += int((ac >= 1.0).sum()) * k
    return tot


def ensemble_errors(a_fields, b_fields):
    errs = {}
    for key in a_fields:
        d = a_fields[key] - b_fields[key]
        errs[key + "_rms"] = float(np.sqrt(np.mean(d ** 2)))
        errs[key + "_max"] = float(np.max(np.abs(d)))
    return errs


ERR_FIELDS = (("eta_n", "ssh_m"), ("U_n", "u_m_s"), ("V_n", "v_m_s"), ("TF", "t_degC"), ("SF", "s_psu"))


MFMA_BF16_PEAK_TFS = 2500.0   # MI355X dense BF16 matrix-core peak


def part_list(model):
    return model.parts if hasattr(model, "parts") else [model]


def probe_parts(model, kernel, steps):
    """Time `kernel` with HIP events over `steps` further (untimed) steps:
    per part, (summed ms, launches)."""
    for p in part_list(model):
        p.probe(kernel)
    model.step(steps)
    model.sync()
    out = []
    for p in part_list(model):
        p.probe(None)
        out.append(p.probe_read())
    return out


def roofline_hbm(kernel, parts, reads, e3, traffic=None):
    """Algorithmic bytes per step of `kernel` over all parts against the summed
    event time of its launches."""
    byts = sum(kernel_bytes(kernel, p.N, p.E, p.D, p.L, p.K, e3) * n for p, (_, n) in zip(parts, reads))
    ms = sum(t for t, _ in reads)
    launches = sum(n for _, n in reads)
    achieved = byts / (ms * 1e-3) / 1e9
    return {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "avg_ms": ms / max(launches, 1),
            "launches": launches, "algorithmic_bytes": byts / max(launches, 1)}


def roofline_mlp(parts, reads):
    """Useful matrix-core work of the learned closure (the network's FLOPs on
    the samples outside the sponge layer) against the summed event time."""
    flop = sum(mlp_samples([p]) * MLP_FLOP_PER_SAMPLE * n for p, (_, n) in zip(parts, reads))
    ms = sum(t for t, _ in reads)
    launches = sum(n for _, n in reads)
    achieved = flop / (ms * 1e-3) / 1e12
    return {"kernel": "k_mlp_closure", "bound": "mfma", "achieved": achieved, "peak": MFMA_BF16_PEAK_TFS,
            "unit": "TFLOP/s", "frac": achieved / MFMA_BF16_PEAK_TFS, "traffic": None, "avg_ms": ms / max(launches, 1),
            "launches": launches, "flop_per_launch": flop / max(launches, 1)}


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["c5", "c2"], default="c5",
                    help="c5: the 64-member ensemble with the learned closure (default); c2: the 1M-node channel")
    ap.add_argument("--scale", default=None, metavar="NX,NY",
                    help="c2 only: mesh nodes along and across the channel instead of the config's")
    ap.add_argument("--probe", default=None, help="kernel timed during the timed region (default: the dominant "
                                                  "one of the path: k_mlp_closure with the learned closure)")
    ap.add_argument("--probe2", default="k_velrhs", help="kernel timed over two further steps (HBM roofline)")
    ap.add_argument("--path", choices=["headline", "mixed-mlp", "fp64", "mixed"], default="headline",
                    help="headline: mixed precision + learned closure, timed beside the plain FP64 drop-in")
    ap.add_argument("--closure-weights", default=os.path.join(ROOT, "fesom-c_amd", "closure_c2.npz"))
    ap.add_argument("--launch-table", default="auto",
                    help="block-size tables from fesomc_amd.autotune (JSON, one per level count); 'auto' = the "
                         "stored tables of this workload and precision if there are some, 'none' = library defaults")
    ap.add_argument("--serial-classes", action="store_true",
                    help="c5 only: finish each level class's step before the next starts (kernel times without "
                         "overlap; the default lets the classes' streams overlap)")
    ap.add_argument("--error-steps", default=None, metavar="N|m2",
                    help="c2 only: continue the headline and the FP64 run in lock step to N internal steps in "
                         "all ('m2': one M2 period, 44 712 s) and report their differences then, over all "
                         "nodes and outside the open-boundary sponge layer")
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
        print(f"bench.py: this build runs one FESOM-C mesh partition (or one ensemble) on one GPU; --gpus "
              f"{args.gpus} is not supported (the multi-GPU halo-exchange design is documented, not built)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__ as ge
    ge.load()
    from fesomc_amd import autotune, closure, gpu, meshgen
    from fesomc_amd.runconfig import RunConfig
    from fesomc_amd.workload import EnsembleSet, build_model

    t_setup = time.perf_counter()
    held = {}
    if args.workload == "c5":
        classes = meshgen.ensemble_classes("c4")
        banks = [gpu.build_meshes(RunConfig(c[1]), c[2]) for c in classes]
    t_mesh = time.perf_counter() - t_setup
    log(f"meshes ready in {t_mesh:.1f} s")
    net = None

    def make(precision, clos):
        nonlocal net
        if clos == gpu.FCG_CLOSURE_MLP and net is None:
            net = closure.load(args.closure_weights)
        if args.workload == "c5":
            return EnsembleSet(classes, precision=precision, closure=clos, banks=banks,
                               net=net if clos == gpu.FCG_CLOSURE_MLP else None, serial=args.serial_classes)
        scale = tuple(int(v) for v in args.scale.split(",")) if args.scale else None
        txt, mesh, ex, model = build_model("c2", scale=scale, precision=precision, closure=clos)
        held["mesh"], held["txt"] = mesh, txt
        if clos == gpu.FCG_CLOSURE_MLP:
            model.attach_mlp(net)
        return model

    def tuned(model, precision):
        table = args.launch_table
        if table == "auto":
            table = autotune.table_path(args.workload, precision)
            table = table if os.path.exists(table) and not args.scale else "none"
        if table != "none":
            autotune.apply_parts(model, autotune.load_parts(table))

    def run(precision, clos, probe):
        t0s = time.perf_counter()
        model = make(precision, clos)
        tuned(model, precision)
        setup = time.perf_counter() - t0s
        log(f"precision {precision} closure {clos}: model ready in {setup:.1f} s")
        model.step(args.warmup)
        model.sync()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        model.step(args.steps)
        model.sync()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        log(f"precision {precision} closure {clos}: {args.steps} steps in {t1 - t0:.3f} s")
        # The roofline kernels are timed over further steps with the level
        # classes stepped in turn: with the classes' streams overlapping, an
        # event pair on one stream also brackets other streams' kernels.
        serial = getattr(model, "serial", None)
        if serial is not None:
            model.serial = True
        reads = probe_parts(model, probe, 2)
        reads2 = probe_parts(model, args.probe2, 2)
        if serial is not None:
            model.serial = serial
        diag = model.diagnostics()
        if diag["status"] != 0 or not math.isfinite(diag["energy"]):
            print(json.dumps({"error": "model state invalid", "diag": diag}), file=sys.stderr)
            sys.exit(1)
        return model, t1 - t0, reads, reads2, setup, diag

    headline = args.path in ("headline", "mixed")   # paths compared against the FP64 drop-in
    prec = 0 if args.path == "fp64" else 1
    clos = gpu.FCG_CLOSURE_MLP if args.path in ("headline", "mixed-mlp") else gpu.FCG_CLOSURE_D3END
    probe = args.probe or ("k_mlp_closure" if clos == gpu.FCG_CLOSURE_MLP else "k_velrhs")
    model, sec, reads, reads2, setup_s, diag = run(prec, clos, probe)
    parts = part_list(model)
    NL = sum(p.N * p.L for p in parts)
    N, E, D = sum(p.N for p in parts), sum(p.E for p in parts), sum(p.D for p in parts)
    nls = NL * args.steps / sec
    e3 = 4 if prec else 8
    # HBM traffic per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes of this workload
    tfile = os.path.join(ROOT, "profiles", f"pmc_traffic_{'mixed' if prec else 'fp64'}_step_{args.workload}.md")
    traffic = (lambda k: pmc_traffic(k, tfile)) if not args.scale else (lambda k: None)
    if probe == "k_mlp_closure":
        roof = roofline_mlp(parts, reads)
        roof["traffic"] = traffic(probe)
    else:
        roof = roofline_hbm(probe, parts, reads, e3, traffic(probe))
    roof2 = roofline_hbm(args.probe2, parts, reads2, e3, traffic(args.probe2))
    desc = {"headline": "mixed precision (FP32 3-D fields, fluxes and tendencies; FP64 elevation/volume, wet/dry, "
                        "BPG) + learned closure (5-128-128-128-2 BF16 MFMA network, d3_end fallback)",
            "fp64": "plain FP64 drop-in", "mixed": "mixed precision, d3_end closure",
            "mixed-mlp": "mixed precision + learned closure (not compared)"}[args.path]
    if args.workload == "c5":
        cls = [{"nsigma": c[0], "members": len(c[2]), "nodes": int(sum(m.nod2D for m in c[2]))} for c in classes]
        workload = (f"c5: 64-member channel ensemble, 50k-400k nodes per member ({N} nodes, {E} elements, {D} "
                    f"edges), 16-32 sigma levels in {len(classes)} classes ({NL} node-levels), Mt {model.par.Mt}, "
                    f"dt {model.par.dt} s, wet/dry, MUSCL tracers, tide amplitude +-10 %, wind +-3 m/s per member; "
                    f"{desc}")
        config = {"workload": workload, "members": model.members, "nodes": N, "elements": E, "node_levels": NL,
                  "classes": cls, "Mt": model.par.Mt, "path": args.path, "serial_classes": args.serial_classes}
    else:
        workload = (f"c2: channel {N} nodes / {E} elements / {D} edges, nsigma {model.K}, Mt {model.par.Mt}, "
                    f"wet/dry, MUSCL tracers; {desc}")
        config = {"workload": workload, "nodes": N, "elements": E, "layers": model.L, "Mt": model.par.Mt,
                  "path": args.path}
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
        "dtype": "f64" if not prec else "f32/f64 (3-D fields and fluxes f32; elevation, wet/dry, BPG f64), mlp bf16",
        "data": "synthetic (seeded channel mesh generator, analytic forcing)",
        "config": config,
        "roofline": roof,
        "roofline_hbm": roof2,
        "setup_s": {"meshes": t_mesh, "model": setup_s},
        "energy": diag["energy"],
    }
    if clos == gpu.FCG_CLOSURE_MLP:
        out["mlp_fallback_column_fraction"] = diag["mlp_fallback"] / (N * (args.warmup + args.steps + 4))
    if headline:
        # the plain FP64 drop-in on the same workload: throughput and the error
        # of the headline path against it after the same steps
        head_fields = {key: model.get(f) for f, key in ERR_FIELDS}
        head_mask = model.get("mask_wd")
        long_run = args.error_steps is not None and args.workload == "c2"
        if not long_run:
            model.close()
        ref, sec64, _, _, setup64, _ = run(0, gpu.FCG_CLOSURE_D3END, "k_velrhs")
        out["fp64_drop_in"] = {"value": NL * args.steps / sec64, "ms_per_step": 1e3 * sec64 / args.steps}
        out["speedup_vs_fp64"] = sec64 / sec
        ref_fields = {key: ref.get(f) for f, key in ERR_FIELDS}
        out["error_vs_fp64_after_steps"] = {"steps": args.warmup + args.steps + 4,
                                            **ensemble_errors(head_fields, ref_fields)}
        out["mask_wd_flips_vs_fp64"] = int((ref.get("mask_wd") != head_mask).sum())
        if long_run:
            done = args.warmup + args.steps + 4
            target = (int(math.ceil(44712.0 / model.par.dt)) if args.error_steps == "m2"
                      else int(args.error_steps))
            while done < target:
                k = min(500, target - done)
                model.step(k)
                ref.step(k)
                done += k
                model.sync()
                ref.sync()
                log(f"lock-step run: {done}/{target} internal steps")
            ac = np.ctypeslib.as_array(model.view.ac, shape=(model.view.nod2D,))
            inner = np.nonzero(ac >= 1.0)[0]
            L = model.L
            hf = {key: model.get(f) for f, key in ERR_FIELDS}
            rf = {key: ref.get(f) for f, key in ERR_FIELDS}
            dm, dr = model.diagnostics(), ref.diagnostics()
            out["error_vs_fp64_long"] = {"steps": done, "seconds": done * model.par.dt,
                                         "status": [dm["status"], dr["status"]], **ensemble_errors(hf, rf)}
            # outside the sponge layer: node fields at inner nodes
            sel = {"ssh_m": inner, "t_degC": (inner[:, None] * L + np.arange(L)).ravel(),
                   "s_psu": (inner[:, None] * L + np.arange(L)).ravel()}
            out["error_vs_fp64_long_outside_sponge"] = ensemble_errors({k: hf[k][v] for k, v in sel.items()},
                                                                       {k: rf[k][v] for k, v in sel.items()})
            out["mask_wd_flips_vs_fp64_long"] = int((ref.get("mask_wd") != model.get("mask_wd")).sum())
            model.close()
        ref.close()
    else:
        model.close()
    if args.workload == "c5":
        for b in banks:
            gpu.free_meshes(b)
    if net is not None:
        net.close()
    if not args.no_cpu_baseline:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
