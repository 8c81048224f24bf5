#!/usr/bin/env python3
"""Autotuner for the construction's launch and layout parameters.

Tuned per (workload kind, mode, periodic, size class = round(log2 n)); the
kinds are i.i.d. uniform seeds, clustered seeds (80 % in 64 Gaussian
clusters of log-uniform sigma, 20 % uniform) and batched builds of many
uneven problems (half uniform, half in 8 clusters, mixed boundaries):

  occupancy16         generators per grid cell x 16 (neighbour gathers and
                      features scan fewer, fuller or more, emptier cells)
  clip_blocks_per_cu  blocks of 4 wavefronts per CU of the persistent
                      one-cell-per-wavefront clip launch
  cells_per_wave      1 (a cell per wavefront, its vertices on the lanes),
                      2 (two cells per wavefront on 32-lane halves) or 64
                      (a cell per lane) for the main lists
  knn_min_tile        smallest LDS tile of the neighbour gather (512, 1024,
                      2048 entries: fewer re-gathers against fewer
                      wavefronts per CU)
  length_order        learned mode: order cells by proposed list length
                      (longest first) or keep grid order (locality)
  first_list_cap      learned mode: the longest first-pass list (the edge of
                      the k buckets the list kernels serve: 64 to 512)

Coordinate descent from the library defaults, each candidate timed as the
minimum over repeated device-resident builds (wall time around a
synchronised build).  The table is written to autotune.json beside this file
and read by `tuned_options`.  Results do not depend on these parameters
(every build is exact); only the time does.

    python -m meshless_voro_amd.autotune [--quick] [--kinds uniform,clustered,batched]

Needs a GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from typing import Dict, Optional

import numpy as np

from . import _native, workloads as W
from .voronoi import MODE_LEARNED, MODE_PLAIN, BuildOptions

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune.json")
SPACE = {
    "occupancy16": [16, 24, 32, 48, 64],
    "clip_blocks_per_cu": [2, 3, 4, 5, 6, 8],
    "cells_per_wave": [1, 2, 64],
    "knn_min_tile": [512, 1024, 2048],
    "length_order": [True, False],
    "first_list_cap": [64, 96, 128, 256, 512],
}
LEARNED_ONLY = ("length_order", "first_list_cap")
DEFAULT = {"occupancy16": 32, "clip_blocks_per_cu": 4, "cells_per_wave": 1,
           "knn_min_tile": 512, "length_order": True, "first_list_cap": 256}
KINDS = ("uniform", "clustered", "batched")


def _key(mode: int, periodic: bool, n: int, kind: str = "uniform") -> str:
    return f"{kind}/{'learned' if mode == MODE_LEARNED else 'plain'}/" \
           f"{int(bool(periodic))}/{int(round(math.log2(max(n, 1))))}"


def _options(mode: int, params: dict, model=None) -> BuildOptions:
    p = dict(DEFAULT)
    p.update(params)
    return BuildOptions(mode=mode, occupancy16=p["occupancy16"],
                        clip_blocks_per_cu=p["clip_blocks_per_cu"],
                        length_order=p["length_order"], cells_per_wave=p["cells_per_wave"],
                        first_list_cap=p["first_list_cap"], knn_min_tile=p["knn_min_tile"],
                        model=model)


def _time_build(L, xt, n, periodic, opt: BuildOptions, reps: int) -> float:
    import torch
    o = opt.to_c()
    a, w = np.zeros(3), np.ones(3)
    best = math.inf
    for r in range(reps + 1):
        h = C.c_void_p()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = L.mv_build_device(C.c_void_p(xt.data_ptr()), n, None, a.ctypes.data, w.ctypes.data,
                               3, int(periodic), C.byref(o), C.byref(h))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if st != 0:
            raise RuntimeError(f"build failed: {st}")
        L.mv_free(h)
        if r > 0:
            best = min(best, dt)
    return best


def _time_batch(L, xt, offsets, per, opt: BuildOptions, reps: int) -> float:
    """Minimum wall time of a batched device-resident build (mv_build_batch_device)."""
    import torch
    o = opt.to_c()
    P = len(per)
    anchors, widths = np.zeros((P, 3)), np.ones((P, 3))
    hs = (C.c_void_p * P)()
    best = math.inf
    for r in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = L.mv_build_batch_device(C.c_void_p(xt.data_ptr()), offsets.ctypes.data, P, None,
                                     anchors.ctypes.data, widths.ctypes.data, per.ctypes.data,
                                     3, C.byref(o), hs)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if st != 0:
            raise RuntimeError(f"batched build failed: {st}")
        for p in range(P):
            L.mv_free(hs[p])
        if r > 0:
            best = min(best, dt)
    return best


def _descend(timer, mode: int, log) -> tuple:
    """Coordinate descent over SPACE with `timer(params) -> seconds`."""
    cur = dict(DEFAULT)
    best_t = timer(cur)
    log(f"  start {cur}: {best_t * 1e3:.2f} ms")
    for name, values in SPACE.items():
        if name in LEARNED_ONLY and mode != MODE_LEARNED:
            continue
        for v in values:
            if v == cur[name]:
                continue
            cand = dict(cur)
            cand[name] = v
            t = timer(cand)
            log(f"  {name}={v}: {t * 1e3:.2f} ms")
            if t < best_t:
                best_t, cur = t, cand
    return cur, best_t


def tune_one(x: np.ndarray, periodic: bool, mode: int, reps: int = 3, log=print) -> dict:
    """Coordinate descent over SPACE for one build; returns the best
    parameters and their time."""
    import torch
    L = _native.lib()
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    n = len(x)
    cur, best_t = _descend(lambda p: _time_build(L, xt, n, periodic, _options(mode, p), reps),
                           mode, log)
    return {"params": cur, "ms": best_t * 1e3, "n": n}


def tune_batch(problems, mode: int, reps: int = 2, log=print) -> dict:
    """Coordinate descent over SPACE for one batched build of
    (positions, periodic) problems in unit boxes."""
    import torch
    L = _native.lib()
    sizes = [len(x) for x, _ in problems]
    offsets = np.zeros(len(problems) + 1, np.uint64)
    offsets[1:] = np.cumsum(sizes)
    per = np.array([int(p) for _, p in problems], np.uint32)
    xt = torch.from_numpy(np.ascontiguousarray(np.concatenate([x for x, _ in problems]))).to(
        "cuda:0")
    cur, best_t = _descend(lambda p: _time_batch(L, xt, offsets, per, _options(mode, p), reps),
                           mode, log)
    return {"params": cur, "ms": best_t * 1e3, "n": int(offsets[-1])}


def load_table(path: str = TABLE) -> Dict[str, dict]:
    try:
        with open(path) as f:
            return json.load(f)["entries"]
    except (OSError, ValueError, KeyError):
        return {}


def tuned_options(n: int, periodic: bool, mode: int = MODE_PLAIN, model=None,
                  table: Optional[Dict[str, dict]] = None,
                  kind: str = "uniform") -> BuildOptions:
    """BuildOptions with the tuned parameters of the nearest size class of the
    same workload kind, mode and boundary kind (library defaults when there
    is none; a batched build passes periodic=True)."""
    table = load_table() if table is None else table
    want = _key(mode, periodic, n, kind)
    prefix = want.rsplit("/", 1)[0] + "/"
    cands = [(abs(int(k.rsplit("/", 1)[1]) - int(want.rsplit("/", 1)[1])), k)
             for k in table if k.startswith(prefix)]
    params = dict(DEFAULT)
    if cands:
        params.update(table[min(cands)[1]]["params"])
    return _options(mode, params, model)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--quick", action="store_true", help="two sizes, two repetitions")
    ap.add_argument("--kinds", default=",".join(KINDS),
                    help="workload kinds to tune (the table keeps the others' entries)")
    args = ap.parse_args()
    kinds = set(args.kinds.split(","))
    if not kinds <= set(KINDS):
        ap.error(f"--kinds: one or more of {', '.join(KINDS)}")
    import torch
    torch.cuda.init()  # torch's HIP runtime comes up before libmvoro's
    reps = 2 if args.quick else 3
    sizes = [1 << 17, 1 << 21] if args.quick else [1 << 15, 1 << 17, 1 << 19, 1 << 21]
    entries = load_table(args.out)
    have_model = True
    try:
        from . import learned
        learned.default_model()
    except Exception:
        have_model = False
    modes = [MODE_PLAIN, MODE_LEARNED] if have_model else [MODE_PLAIN]
    log = lambda s: print(s, flush=True)  # noqa: E731
    # held-out seeds (streams the benchmark sets do not use)
    for n in sizes if "uniform" in kinds else []:
        for periodic in (True, False):
            x = W.uniform(n, 7000 + n.bit_length() + int(periodic))
            for mode in modes:
                k = _key(mode, periodic, n)
                print(f"tune {k}", flush=True)
                entries[k] = tune_one(x, periodic, mode, reps, log=log)
                print(f"best {k}: {entries[k]}", flush=True)
    cl_sizes = [1 << 20] if args.quick else [1 << 20, 1 << 22]
    for n in cl_sizes if "clustered" in kinds else []:
        x = W.clustered(n, 7100 + n.bit_length())
        for mode in modes:
            k = _key(mode, True, n, "clustered")
            print(f"tune {k}", flush=True)
            entries[k] = tune_one(x, True, mode, 2, log=log)
            print(f"best {k}: {entries[k]}", flush=True)
    probs = W.batched_problems(128 if args.quick else 512, 7200)
    for mode in modes if "batched" in kinds else []:
        k = _key(mode, True, sum(len(x) for x, _ in probs), "batched")
        print(f"tune {k}", flush=True)
        entries[k] = tune_batch(probs, mode, 2, log=log)
        print(f"best {k}: {entries[k]}", flush=True)
    with open(args.out, "w") as f:
        json.dump({"space": SPACE, "default": DEFAULT, "entries": entries}, f, indent=1,
                  sort_keys=True)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
