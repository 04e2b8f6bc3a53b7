"""Learned initial guess for the p' solve: training data from the MSI
solver, training of the U-Net (mach2d.unet) on the GPU kernels, and the
guess-then-MSI solve with plain MSI from zero as the fallback.

    python -m mach2d.learned train --samples 20000 --epochs 4 --out unet_p.m2dw
    python -m mach2d.learned eval --weights unet_p.m2dw

Training pairs: p' systems on 258x130 ogive grids with seeded wall
stretching (q drawn in [1.02, 1.12]) and right-hand sides that mix smooth
random modes with white noise; each target is the system's FP64 MSI solution
to a 1e-10 residual drop, solved in batches by the batched solver.

The network predicts the shape of p' (targets are scaled to unit RMS); the
guessed solve takes the amplitude that minimises the energy functional of A,
alpha = (g.b)/(g.A g) (m2d_guess_scale_energy): it weighs the smooth error
components, the ones MSI removes slowest, where the residual-minimising
amplitude (m2d_guess_scale) is dominated by the oscillatory part of the
residual and shrinks a smooth guess towards zero.  A guess with g.A g <= 0
gets alpha = 0, i.e. the zero guess.  Plain MSI from the zero guess is the
fallback and the judge: a scaled guess whose residual norm is not below
||b|| (the zero guess's residual) is replaced by the zero guess.  The solve
stops when ||r|| < gamma ||b||, the zero guess's threshold, so a guess is
never asked for more accuracy than the zero-guess solve delivers.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import problems as P

TRAIN_NX, TRAIN_NY = 258, 130


def training_systems(count, seed=0):
    """Seeded training systems: (mats {q: a (9N)}, qs (count), b (count, N)).
    Stretching q takes 11 values in [1.02, 1.12]; 60% smooth right-hand
    sides, 40% white noise."""
    rng = np.random.default_rng(seed)
    qs = np.round(rng.uniform(1.02, 1.12, size=count), 2)
    smooth = rng.uniform(size=count) < 0.6
    mats = {}
    b = np.empty((count, TRAIN_NX * TRAIN_NY))
    for i in range(count):
        q = float(qs[i])
        if q not in mats:
            mats[q] = P.pcorr_system(TRAIN_NX, TRAIN_NY, seed=0, q=q).a.reshape(-1)
        b[i] = P.rhs_field(TRAIN_NX, TRAIN_NY, "smooth" if smooth[i] else "gauss",
                           np.random.default_rng((seed, i)))
    return mats, qs, b


def solve_targets(mats, qs, b, batch=256, log=print):
    """FP64 MSI solutions (1e-10 drop) of the systems, batch by batch."""
    import torch

    from .solvers import BatchMSI
    count = b.shape[0]
    x = np.empty_like(b)
    dm = {q: torch.tensor(a, device="cuda") for q, a in mats.items()}
    for s0 in range(0, count, batch):
        n = min(batch, count - s0)
        bm = BatchMSI([(TRAIN_NX, TRAIN_NY)] * n, precision="fp64")
        ta = [dm[float(qs[s0 + i])] for i in range(n)]
        tb = [torch.tensor(b[s0 + i], device="cuda") for i in range(n)]
        tv = [torch.zeros(b.shape[1], dtype=torch.float64, device="cuda") for _ in range(n)]
        bm.factor(ta)
        rc, st = bm.solve(ta, tb, tv, 1e-10, 20000)
        if rc or any(y.status for y in st):
            raise RuntimeError("training solve failed: %s" % rc)
        for i in range(n):
            x[s0 + i] = tv[i].cpu().numpy()
        bm.close()
        log("targets %d / %d" % (s0 + n, count))
    return x


def make_dataset(net, mats, qs, b, x):
    """Network inputs (bf16, on the device), targets x / s and masks."""
    import torch
    n = b.shape[0]
    h, w = net.padded(TRAIN_NX, TRAIN_NY)
    X = torch.empty(n, h, w, 8, dtype=torch.bfloat16, device="cuda")
    T = torch.zeros(n, h, w, dtype=torch.float32, device="cuda")
    M = torch.zeros(n, h, w, dtype=torch.float32, device="cuda")
    M[:, :TRAIN_NY, :TRAIN_NX] = 1.0
    dm = {q: torch.tensor(a, device="cuda") for q, a in mats.items()}
    for i in range(n):
        a = mats[float(qs[i])]
        ap = a.reshape(9, -1)[4]
        qv = np.where(ap != 0, b[i] / np.where(ap != 0, ap, 1.0), 0.0)
        s = float(np.sqrt(np.mean(qv * qv))) or 1.0
        X[i:i + 1] = net.pack_input(dm[float(qs[i])], torch.tensor(b[i], device="cuda"),
                                    TRAIN_NX, TRAIN_NY, s)
        t = x[i] / s
        t = t / (float(np.sqrt(np.mean(t * t))) or 1.0)   # unit RMS: the network learns the shape
        T[i, :TRAIN_NY, :TRAIN_NX] = torch.tensor(t.reshape(TRAIN_NY, TRAIN_NX),
                                                  dtype=torch.float32, device="cuda")
    return X, T, M


def train(net, X, T, M, epochs=4, batch=16, lr=1e-3, seed=0, log=print):
    import torch
    n = X.shape[0]
    g = torch.Generator().manual_seed(seed)
    steps = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to("cuda")
        tot, cnt = 0.0, 0
        t0 = time.time()
        for s0 in range(0, n - batch + 1, batch):
            idx = perm[s0:s0 + batch]
            loss = net.loss_and_backward(X[idx].contiguous(), T[idx].contiguous(),
                                         M[idx].contiguous())
            net.adam_step(lr=lr * (0.5 ** ep))
            tot += float(loss.item())
            cnt += 1
            steps += 1
            if steps % 200 == 0:
                log("step %d: loss %.4g" % (steps, float(loss.item())))
        log("epoch %d: mean loss %.4g over %d steps (%.1fs)" % (ep, tot / max(cnt, 1), cnt,
                                                                time.time() - t0))
    return steps


def judge_guess(dev, a, rhs, var, rg=None):
    """The residual-norm judge of a guess held in `var`: plain MSI from the
    zero guess is the fallback, so the guess is kept only when its residual
    ||A var - rhs|| is below the zero guess's, ||rhs||; otherwise `var` is
    zeroed.  Returns (used, ||A var - rhs||, ||rhs||) for what `var` holds
    afterwards.  `rg` passes a residual norm already computed for `var`."""
    import torch
    rb = float(torch.linalg.norm(rhs))
    if rg is None:
        rg = dev.residual_norm(a, rhs, var)
    if not (np.isfinite(rg) and rg < rb):
        var.zero_()
        return False, rb, rb
    return True, rg, rb


def prepare_guess(dev, net, a, rhs, var, nx, ny):
    """The network's guess in `var`, scaled by the energy-minimising
    amplitude (handle `dev`, FP64 residual norms), then judged by
    judge_guess: kept only if its residual is below ||rhs||, else the zero
    guess.  Returns (used, ||A var - rhs||, ||rhs||)."""
    import torch
    net.guess(a, rhs, nx, ny, var)
    alpha = dev.guess_scale(a, rhs, var, energy=True)
    if not (np.isfinite(alpha) and alpha != 0.0):
        var.zero_()
        rb = float(torch.linalg.norm(rhs))
        return False, rb, rb
    return judge_guess(dev, a, rhs, var)


def prepare_guesses(dev, net, a_list, rhs_list, var_list, nx, ny):
    """prepare_guess for several systems of one grid size, with one network
    pass over all of them and the same residual-norm judge per system.
    Returns [(used, ||A var - rhs||, ||rhs||)]."""
    import torch
    net.guess_many(a_list, rhs_list, nx, ny, var_list)
    rbs = [float(v) for v in torch.stack([torch.linalg.norm(r) for r in rhs_list]).cpu()]
    alphas, rnorms = dev.guess_prepare_many(a_list, rhs_list, var_list, energy=True)
    res = []
    for var, rb, alpha, rg in zip(var_list, rbs, alphas, rnorms):
        if not (np.isfinite(alpha) and alpha != 0.0 and np.isfinite(rg) and rg < rb):
            var.zero_()
            res.append((False, rb, rb))
        else:
            res.append((True, rg, rb))
    return res


def guessed_solve(dev, net, a, rhs, var, nx, ny, gamma=1e-10, nitm=20000):
    """MSI (handle `dev`, factored) from the network's guess, stopping at
    gamma ||rhs||.  Returns (status, used_guess, ||r(guess)||, ||rhs||)."""
    used, rg, rb = prepare_guess(dev, net, a, rhs, var, nx, ny)
    st = dev.solve(a, rhs, var, gamma * rb / rg, nitm)
    return st, used, rg, rb


def default_weights():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_p.m2dw")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=default_weights())
    ap.add_argument("--weights", default=default_weights())
    args = ap.parse_args(argv)

    def log(m):
        print("[learned] " + m, file=sys.stderr, flush=True)

    from .unet import UNet
    if args.cmd == "train":
        t0 = time.time()
        mats, qs, b = training_systems(args.samples, args.seed)
        log("systems: %d in %.1fs" % (args.samples, time.time() - t0))
        x = solve_targets(mats, qs, b, log=log)
        log("targets solved at %.1fs" % (time.time() - t0))
        net = UNet(seed=args.seed)
        X, T, M = make_dataset(net, mats, qs, b, x)
        del b, x
        steps = train(net, X, T, M, args.epochs, args.batch, args.lr, args.seed, log)
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        net.save(args.out)
        log("saved %s after %d steps, %.1fs" % (args.out, steps, time.time() - t0))
        return 0
    import torch

    from .solvers import DeviceMSI
    net = UNet()
    net.load(args.weights)
    s = P.pcorr_system(**{k: v for k, v in P.CFG3.items() if k not in ("gamma",)})
    a = torch.tensor(s.a.reshape(-1), device="cuda")
    rhs = torch.tensor(s.rhs, device="cuda")
    out = {}
    for prec in ("fp64", "fp32_ir"):
        dev = DeviceMSI(s.nx, s.ny, prec)
        dev.factor(a)
        v0 = torch.zeros(s.N, dtype=torch.float64, device="cuda")
        st0 = dev.solve(a, rhs, v0, 1e-10, 20000)
        v1 = torch.zeros_like(v0)
        st1, used, rg, rb = guessed_solve(dev, net, a, rhs, v1, s.nx, s.ny)
        out[prec] = dict(zero_nit=st0.nit, guess_nit=st1.nit, used=used, r_guess_over_b=rg / rb,
                         rel_diff=float(torch.linalg.norm(v1 - v0) / torch.linalg.norm(v0)))
        dev.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
