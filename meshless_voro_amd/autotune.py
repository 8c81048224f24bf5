#!/usr/bin/env python3
"""Autotuner for the construction's launch and layout parameters.

Tuned per (mode, periodic, size class = round(log2 n)):

  occupancy16         generators per grid cell x 16 (neighbour gathers and
                      features scan fewer, fuller or more, emptier cells)
  clip_blocks_per_cu  blocks of 4 wavefronts per CU of the persistent
                      one-cell-per-wavefront clip launch
  length_order        learned mode: order cells by proposed list length
                      (longest first) or keep grid order (locality)

Coordinate descent from the library defaults, each candidate timed as the
minimum over repeated device-resident builds (HIP events inside the library
give the same ranking; wall time around a synchronised build is used).  The
table is written to autotune.json beside this file and read by
`tuned_options`.  Results do not depend on these parameters (every build is
exact); only the time does.

    python -m meshless_voro_amd.autotune [--quick]

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
    "clip_blocks_per_cu": [2, 3, 4, 6, 8],
    "length_order": [True, False],
}
DEFAULT = {"occupancy16": 32, "clip_blocks_per_cu": 4, "length_order": True}


def _key(mode: int, periodic: bool, n: int) -> str:
    return f"{'learned' if mode == MODE_LEARNED else 'plain'}/{int(bool(periodic))}/" \
           f"{int(round(math.log2(max(n, 1))))}"


def _options(mode: int, params: dict, model=None) -> BuildOptions:
    return BuildOptions(mode=mode, occupancy16=params["occupancy16"],
                        clip_blocks_per_cu=params["clip_blocks_per_cu"],
                        length_order=params["length_order"], model=model)


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


def tune_one(x: np.ndarray, periodic: bool, mode: int, reps: int = 3, log=print) -> dict:
    """Coordinate descent over SPACE for one workload; returns the best
    parameters and their time."""
    import torch
    L = _native.lib()
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    n = len(x)
    cur = dict(DEFAULT)
    best_t = _time_build(L, xt, n, periodic, _options(mode, cur), reps)
    log(f"  start {cur}: {best_t * 1e3:.2f} ms")
    for name, values in SPACE.items():
        if name == "length_order" and mode != MODE_LEARNED:
            continue
        for v in values:
            if v == cur[name]:
                continue
            cand = dict(cur)
            cand[name] = v
            t = _time_build(L, xt, n, periodic, _options(mode, cand), reps)
            log(f"  {name}={v}: {t * 1e3:.2f} ms")
            if t < best_t:
                best_t, cur = t, cand
    return {"params": cur, "ms": best_t * 1e3, "n": n}


def load_table(path: str = TABLE) -> Dict[str, dict]:
    try:
        with open(path) as f:
            return json.load(f)["entries"]
    except (OSError, ValueError, KeyError):
        return {}


def tuned_options(n: int, periodic: bool, mode: int = MODE_PLAIN, model=None,
                  table: Optional[Dict[str, dict]] = None) -> BuildOptions:
    """BuildOptions with the tuned parameters of the nearest size class of the
    same mode and boundary kind (library defaults when there is none)."""
    table = load_table() if table is None else table
    want = _key(mode, periodic, n)
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
    args = ap.parse_args()
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
    for n in sizes:
        for periodic in (True, False):
            x = W.uniform(n, 7000 + n.bit_length() + int(periodic))
            for mode in ([MODE_PLAIN, MODE_LEARNED] if have_model else [MODE_PLAIN]):
                k = _key(mode, periodic, n)
                print(f"tune {k}", flush=True)
                entries[k] = tune_one(x, periodic, mode, reps,
                                      log=lambda s: print(s, flush=True))
                print(f"best {k}: {entries[k]}", flush=True)
    with open(args.out, "w") as f:
        json.dump({"space": SPACE, "default": DEFAULT, "entries": entries}, f, indent=1,
                  sort_keys=True)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
