#!/usr/bin/env python3
"""Train the learned neighbour-count proposal and write its checkpoint.

Training sets are synthetic seed sets built with the plain (exact) method,
whose per-cell records give the targets: the list length each cell needed
and its final security radius.  The held-out calibration set (uniform seeds
from a stream no training set uses) picks the multipliers on the proposed
list length (kscale) and gather radius (rscale) by timing learned builds.

    python -m meshless_voro_amd.train_proposal --out learned_k.mvlk

Needs a GPU (training and inference run in libmvoro).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import time

import numpy as np

from . import _native, learned, workloads as W
from .voronoi import MODE_LEARNED, MODE_PLAIN, BuildOptions


def training_sets(scale: int):
    """(name, seeds, periodic) of the training mix."""
    n = scale
    yield "uniform_periodic", W.uniform(n, 1001), True
    yield "uniform_periodic_b", W.uniform(n, 1002), True
    yield "uniform_open", W.uniform(n, 1003), False
    yield "clustered_periodic", W.clustered(n, 1004), True
    yield "clustered_open", W.clustered(n // 2, 1005), False
    c = int(round(n ** (1 / 3)))
    yield "perturbed_grid", W.perturbed_grid(c, 0.2, 1006), True


def device_build(xt, n, periodic, opt):
    L = _native.lib()
    a = np.zeros(3)
    w = np.ones(3)
    o = opt.to_c()
    h = C.c_void_p()
    st = L.mv_build_device(C.c_void_p(xt.data_ptr()), n, None, a.ctypes.data, w.ctypes.data,
                           3, int(periodic), C.byref(o), C.byref(h))
    if st != 0:
        raise RuntimeError(f"build failed: {st}")
    s = _native.mv_stats()
    L.mv_get_stats(h, C.byref(s))
    L.mv_free(h)
    return {k: getattr(s, k) for k, _ in _native.mv_stats._fields_}


def time_mode(x, periodic, opt, reps=3):
    import torch
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    n = len(x)
    device_build(xt, n, periodic, opt)
    torch.cuda.synchronize()
    best = 1e30
    st = None
    for _ in range(reps):
        t0 = time.perf_counter()
        st = device_build(xt, n, periodic, opt)
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best, st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=learned.CHECKPOINT)
    ap.add_argument("--scale", type=int, default=1 << 20)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--tau", type=float, default=0.95)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--calib-n", type=int, default=1 << 20)
    args = ap.parse_args()
    # torch (device buffers of the timed builds) brings its own HIP runtime,
    # which must initialise before libmvoro's
    import torch
    torch.cuda.init()

    Fs, Ts = [], []
    for name, x, periodic in training_sets(args.scale):
        t0 = time.perf_counter()
        from .voronoi import Voronoi
        v = Voronoi.build(x, [0, 0, 0], [1, 1, 1], 3, periodic)
        F, h = learned.features(x, [0, 0, 0], [1, 1, 1], 3, periodic)
        T = learned.targets(v, h)
        Fs.append(F)
        Ts.append(T)
        k = np.exp(T[:, 0])
        print(f"set {name}: n={len(x)} k* mean {k.mean():.1f} p95 {np.quantile(k, .95):.0f} "
              f"max {k.max():.0f} ({time.perf_counter() - t0:.1f} s)", flush=True)
    F = np.concatenate(Fs)
    T = np.concatenate(Ts)
    t0 = time.perf_counter()
    model, loss = learned.train(F, T, steps=args.steps, batch=args.batch, seed=1, lr=args.lr,
                                tau=args.tau)
    print(f"trained {args.steps} steps x {args.batch} rows on {len(F)} rows: loss {loss:.4f} "
          f"({time.perf_counter() - t0:.1f} s)", flush=True)
    Y = model.predict(F)
    print("coverage of the tau-quantile per output:",
          [float(np.mean(T[:, c] <= Y[:, c])) for c in range(2)], flush=True)

    # calibration on held-out seeds
    xc = W.uniform(args.calib_n, 2001)
    t_plain, st_plain = time_mode(xc, True, BuildOptions(mode=MODE_PLAIN))
    print(f"calib plain: {t_plain * 1e3:.1f} ms, redone {st_plain['cells_redone']}", flush=True)
    best = None
    for rscale in (0.0, 1.05, 1.1, 1.15):
        for kscale in (0.9, 1.0, 1.1, 1.2, 1.3):
            model.kscale = kscale
            model.rscale = rscale
            t, st = time_mode(xc, True, BuildOptions(mode=MODE_LEARNED, model=model))
            print(f"calib kscale {kscale} rscale {rscale}: {t * 1e3:.1f} ms "
                  f"(knn {st['ms_knn']:.1f} clip {st['ms_clip']:.1f}), redone "
                  f"{st['cells_redone']}, passes {st['passes']}", flush=True)
            if best is None or t < best[0]:
                best = (t, kscale, rscale)
    model.kscale, model.rscale = best[1], best[2]
    model.save(args.out)
    print(json.dumps({"out": args.out, "kscale": best[1], "rscale": best[2],
                      "calib_ms_learned": best[0] * 1e3, "calib_ms_plain": t_plain * 1e3,
                      "loss": loss}), flush=True)


if __name__ == "__main__":
    main()
