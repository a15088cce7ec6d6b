"""Training and storage of the learned vertical closure (fcg_mlp_* entry points).

The network is trained on columns the plain model generates: the model runs
with d3_end (src/fv_mixing.F90:73-347) and, right after each vertical-mixing
call, ``fcg_mlp_collect`` writes the features of every level of every
``stride``-th column and d3_end's log10 snu / log10 Kv as targets.  Training
steps run on the device (``fcg_mlp_train_step``: BF16 matrix-core forward and
backward, FP32 accumulation, Adam).  PyTorch only provides device buffers and
random minibatch indices.

    python -c "import __graft_entry__ as g; g.load(); from fesomc_amd import closure; closure.main()" \
        --config c2 --steps 24 --every 532 --stride 12 --iters 60000 --out closure_c2.npz

collects 24 steps spread over one M2 cycle of the 1M-node channel (every
532nd internal step), every 12th column: about 2.0 M columns (48 M samples).
"""
from __future__ import annotations

import argparse
import math
import time

import numpy as np

from . import gpu


def collect(model, steps, stride=1, torch=None, every=1, log=None):
    """Advance `model` (plain closure) `steps` internal steps, collecting
    (features, targets) after every vertical-mixing call; with every > 1 the
    model advances every-1 further steps before each collected one.  Returns
    device tensors."""
    import torch as _t
    torch = torch or _t
    cols = (model.N + stride - 1) // stride
    cap = cols * model.K
    Xs, Ts = [], []
    for i in range(steps):
        if every > 1:
            model.step(every - 1)
        for ph in gpu.Model.PHASES:
            model.phase(ph)
            if ph == "fcg_vertical_mixing":
                X = torch.empty((cap, gpu.MLP_FEATURES), dtype=torch.float32, device="cuda")
                T = torch.empty((cap, 2), dtype=torch.float32, device="cuda")
                n = model.collect(X.data_ptr(), T.data_ptr(), cap, stride)
                Xs.append(X[:n])
                Ts.append(T[:n])
        if log:
            log(f"collected step {i + 1}/{steps}")
    return torch.cat(Xs), torch.cat(Ts)


def fit(X, T, iters=2000, batch=16384, lr=3e-3, seed=0, margin=0.5, log=None):
    """Normalise, train, and return a gpu.Mlp whose accepted feature box is
    the training range widened by `margin` standard deviations."""
    import torch
    xm, xs = X.mean(0), X.std(0).clamp_min(1e-6)
    ym, ys = T.mean(0), T.std(0).clamp_min(1e-6)
    Xn, Tn = ((X - xm) / xs).contiguous(), ((T - ym) / ys).contiguous()
    lo, hi = Xn.min(0).values - margin, Xn.max(0).values + margin
    net = gpu.Mlp(seed=seed)
    net.set_norm(xm.cpu().numpy(), xs.cpu().numpy(), ym.cpu().numpy(), ys.cpu().numpy(), lo.cpu().numpy(),
                 hi.cpu().numpy())
    g = torch.Generator(device="cuda").manual_seed(seed)
    n = Xn.shape[0]
    for it in range(iters):
        idx = torch.randint(0, n, (batch,), device="cuda", generator=g)
        xb, tb = Xn[idx].contiguous(), Tn[idx].contiguous()
        torch.cuda.synchronize()
        step_lr = lr * 0.5 * (1.0 + math.cos(math.pi * it / iters)) + 1e-2 * lr
        net.train_step(xb.data_ptr(), tb.data_ptr(), batch, step_lr)
        if log and (it % 500 == 0 or it == iters - 1):
            log(f"iter {it}")
    torch.cuda.synchronize()
    return net


def validation_loss(net, X, T, n_max=1 << 20):
    """0.5*mean |f(x)-t|^2 in normalised units on (a prefix of) X, T."""
    import torch
    xm, xs, ym, ys = (torch.tensor(a, device="cuda") for a in net.norm[:4])
    n = min(X.shape[0], n_max)
    Xn = ((X[:n] - xm) / xs).contiguous()
    Tn = ((T[:n] - ym) / ys).contiguous()
    Y = torch.empty_like(Tn)
    torch.cuda.synchronize()
    net.forward(Xn.data_ptr(), Y.data_ptr(), n)
    return float(0.5 * ((Y - Tn) ** 2).sum(1).mean())


def save(net, path):
    xm, xs, ym, ys, lo, hi = net.norm
    np.savez(path, params=net.params(), x_mean=xm, x_std=xs, y_mean=ym, y_std=ys, x_lo=lo, x_hi=hi)


def load(path):
    d = np.load(path, allow_pickle=False)
    net = gpu.Mlp(seed=0)
    net.set_params(d["params"])
    net.set_norm(d["x_mean"], d["x_std"], d["y_mean"], d["y_std"], d["x_lo"], d["x_hi"])
    return net


def main():
    ap = argparse.ArgumentParser(description="train the learned closure on columns of the plain model")
    ap.add_argument("--config", default="c2")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--stride", type=int, default=10)
    ap.add_argument("--every", type=int, default=1)
    ap.add_argument("--iters", type=int, default=60000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    from .workload import build_model
    t0 = time.time()
    _, _, _, model = build_model(args.config)
    X, T = collect(model, args.steps, args.stride, every=args.every, log=lambda m: print(m, flush=True))
    d = model.diagnostics()
    if d["status"] != 0 or not math.isfinite(d["energy"]):
        raise SystemExit(f"the plain model failed while collecting: {d}")
    model.close()
    print(f"collected {X.shape[0]} samples in {time.time() - t0:.1f} s", flush=True)
    n = X.shape[0]
    perm = np.random.default_rng(0).permutation(n)
    import torch
    perm = torch.tensor(perm, device="cuda")
    ntr = int(0.9 * n)
    Xtr, Ttr, Xva, Tva = X[perm[:ntr]], T[perm[:ntr]], X[perm[ntr:]], T[perm[ntr:]]
    net = fit(Xtr, Ttr, iters=args.iters, batch=args.batch, lr=args.lr, log=lambda m: print(m, flush=True))
    print(f"validation loss {validation_loss(net, Xva, Tva):.4e} (normalised units), "
          f"train loss {validation_loss(net, Xtr, Ttr):.4e}, {time.time() - t0:.1f} s", flush=True)
    save(net, args.out)


if __name__ == "__main__":
    main()
