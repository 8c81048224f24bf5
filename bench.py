#!/usr/bin/env python3
"""Headline benchmark: 3D Voronoi cells/s at 2,097,152 seeds.

Workload: 2,097,152 (128^3) seeds drawn i.i.d. uniform in the unit cube from
SplitMix64 stream 2 (meshless_voro_amd.workloads.uniform_2m), periodic box,
FP64 positions, volume/centroid/face-area integrals and cell->face links.
One step = one complete Voronoi::build (grid, neighbour lists, clipping,
integrals, faces, links) from device-resident seeds to device-resident
results.  `value` is the cells/s of the fastest construction mode; the plain
GPU form (k = 27 per pass, doubled for the cells whose security radius was
not reached) is reported beside it, with their connectivity checked equal.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
Prints one JSON line on stdout.

Roofline of the dominant kernel: its average launch duration comes from HIP
events the library records around each launch on the build's stream
(mv_stats), over the timed builds; the algorithmic work per launch is the
per-unit figure (FP64 flops per completed cell for the clip kernel; bytes
per list for the neighbour-list kernel) times the units one launch
processes.  `traffic` is the HBM bytes per launch measured with rocprofv3
FETCH_SIZE / WRITE_SIZE (scaled by the same counters on a plain copy of
known size, see --calibrate-copy), read from profiles/traffic_<kernel>.md
when that file exists.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D Voronoi cells/s at 2.1M seeds: learned-k+batched vs plain GPU, same topology"
# algorithmic FP64 operations (add/sub/mul/div/sqrt, each 1) per cell of
# this workload, counted by the oracle's instrumented build on 200,000 seeds
# of the same distribution (tests/test_oracle.py::test_flop_count_per_uniform_cell
# pins the range)
FLOPS_PER_CELL = 28786.0
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2097152)
    ap.add_argument("--modes", default="auto")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=2097152)
    ap.add_argument("--batched", action="store_true",
                    help="also measure many problems in one build (512 problems, ~10.6 M "
                         "seeds; the clustered problems take tens of seconds)")
    ap.add_argument("--clustered", action="store_true",
                    help="also time 4,194,304 clustered periodic seeds (80 %% in 64 Gaussian "
                         "clusters, sigma log-uniform in [1e-3, 1e-1]) in both modes")
    ap.add_argument("--host", action="store_true",
                    help="also time mv_build with host seeds in and host results out "
                         "(upload, build, download of every output array)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (elementwise kernel) so that "
                         "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE can be scaled")
    return ap.parse_args()


def calibrate_copy():
    """A plain copy of known size for scaling the HBM counters: y = x + 0
    over 2^27 doubles (1 GiB read, 1 GiB written) in torch's elementwise
    kernel."""
    import torch
    x = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    torch.add(x, 0.0, out=y)
    torch.cuda.synchronize()
    del x, y


def batched_c5(L, mode, reps=3, only_uniform=False):
    """512 independent problems (10^3..10^5 seeds each, ~10.6 M in total; even
    problems uniform, odd problems 8 wrapped Gaussian clusters; periodic for
    p % 4 < 2), device-resident: one batched build (mv_build_batch_device)
    against one build per problem (mv_build_device), full and with a 10 %
    mask on every problem (build_partial).  only_uniform keeps the 256
    uniform problems."""
    import numpy as np
    import torch
    from meshless_voro_amd import _native, workloads as W
    from meshless_voro_amd.voronoi import BuildOptions, VoronoiError
    probs = W.batched_problems(512, 5)
    if only_uniform:
        probs = probs[0::2]
    P = len(probs)
    sizes = [len(x) for x, _ in probs]
    offsets = np.zeros(P + 1, np.uint64)
    offsets[1:] = np.cumsum(sizes)
    xyz = np.ascontiguousarray(np.concatenate([x for x, _ in probs]))
    per = np.array([int(p) for _, p in probs], np.uint32)
    anchors = np.zeros((P, 3))
    widths = np.ones((P, 3))
    mask = np.ascontiguousarray(np.concatenate(
        [W.partial_mask(n, 0.1, 1000 + i) for i, n in enumerate(sizes)]).astype(np.uint8))
    xt = torch.from_numpy(xyz).to("cuda:0")
    mt = torch.from_numpy(mask).to("cuda:0")
    from meshless_voro_amd.autotune import tuned_options
    bo = tuned_options(int(sum(sizes)), True, mode, kind="batched")
    opt = bo.to_c()
    hs = (C.c_void_p * P)()

    def batched(use_mask):
        st = L.mv_build_batch_device(C.c_void_p(xt.data_ptr()), offsets.ctypes.data, P,
                                     C.c_void_p(mt.data_ptr()) if use_mask else None,
                                     anchors.ctypes.data, widths.ctypes.data, per.ctypes.data, 3,
                                     C.byref(opt), hs)
        if st != 0:
            raise VoronoiError(st)
        for p in range(P):
            L.mv_free(hs[p])

    def sequential(use_mask):
        a, w = np.zeros(3), np.ones(3)
        for p in range(P):
            if p % 32 == 31:
                print(f"bench:     problem {p + 1}/{P}", file=sys.stderr, flush=True)
            o = int(offsets[p])
            h = C.c_void_p()
            st = L.mv_build_device(C.c_void_p(xt.data_ptr() + 24 * o), sizes[p],
                                   C.c_void_p(mt.data_ptr() + o) if use_mask else None,
                                   a.ctypes.data, w.ctypes.data, 3, int(per[p]), C.byref(opt),
                                   C.byref(h))
            if st != 0:
                raise VoronoiError(st)
            L.mv_free(h)

    def timed(fn, *a):
        print(f"bench:   {fn.__name__} mask={a[0]}", file=sys.stderr, flush=True)
        if reps > 1:
            fn(*a)
            torch.cuda.synchronize()
        best = 1e30
        for _ in range(reps):
            t0 = time.perf_counter()
            fn(*a)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    def captured(use_mask, reps_g):
        """The same batched build recorded once as a hipGraph
        (mv_graph_create) and replayed (mv_graph_launch): creation time,
        median replay time, and whether replays ran the recording (also after
        every seed moved by ~1e-4)."""
        from meshless_voro_amd import BuildGraph
        prob_list = [(int(offsets[p]), [0., 0., 0.], [1., 1., 1.], bool(per[p])) for p in range(P)]
        t0 = time.perf_counter()
        g = BuildGraph(xt, prob_list, mask=mt if use_mask else None, options=bo)
        t_create = time.perf_counter() - t0
        flags = []

        def replay():
            rep = C.c_uint32(0)
            st = L.mv_graph_launch(g._h, hs, C.byref(rep))
            if st != 0:
                raise VoronoiError(st)
            flags.append(int(rep.value))
            for p in range(P):
                L.mv_free(hs[p])

        replay()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps_g):
            t0 = time.perf_counter()
            replay()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        keep = xt.clone()
        gen = torch.Generator(device="cuda:0")
        gen.manual_seed(7)
        xt.add_(1e-4 * torch.randn(xt.shape, generator=gen, device="cuda:0", dtype=xt.dtype))
        xt.remainder_(1.0)
        replay()
        moved_ok = flags[-1]
        moved_viol = g.last_violation()
        xt.copy_(keep)
        torch.cuda.synchronize()
        g.close()
        return {"graph_create_ms": 1e3 * t_create, "graph_ms": 1e3 * float(np.median(ts)),
                "graph_replayed": sum(flags[:-1]) / max(len(flags) - 1, 1),
                "graph_replayed_after_move": moved_ok, "graph_violation_after_move": moved_viol}

    n = int(offsets[-1])
    nm = int(mask.sum())
    out = {"problems": P, "seeds": n, "periodic_problems": int(per.sum())}
    for name, use_mask, cells in (("full", False, n), ("partial_10pct", True, nm)):
        tb = timed(batched, use_mask)
        ts = timed(sequential, use_mask)
        out[name] = {"cells": cells, "batched_ms": 1e3 * tb, "sequential_ms": 1e3 * ts,
                     "batched_cells_per_s": cells / tb, "sequential_cells_per_s": cells / ts,
                     "batched_speedup": ts / tb}
        print(f"bench:   captured {name}", file=sys.stderr, flush=True)
        out[name].update(captured(use_mask, max(reps, 3)))
        print(f"bench:   {name}: {json.dumps(out[name])}", file=sys.stderr, flush=True)
    return out


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.md")
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt, re.M)
    return float(m.group(1)) if m else None


def clustered_4m(L, modes, reps=2):
    """4,194,304 clustered periodic seeds: device time per full build (HIP
    events, after one warm-up build) in each mode, and whether both modes
    give the same connectivity."""
    import numpy as np
    import torch
    from meshless_voro_amd import workloads as W
    x = W.clustered_4m()
    n = len(x)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    res, dig = {}, {}
    for name, mode in modes:
        print(f"bench: clustered 4M, mode {name}", file=sys.stderr, flush=True)
        h, _ = build_device(L, xt, n, mode, kind="clustered")
        dig[name] = connectivity_digest(L, h)
        L.mv_free(h)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            h, st = build_device(L, xt, n, mode, kind="clustered")
            L.mv_free(h)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms_per_build": ms, "cells_per_s": n / (ms / 1e3),
                     "passes": st["passes"], "big_cells": st["big_cells"],
                     "k_last": st["k_last"]}
    if "learned" in dig and "plain" in dig:
        res["same_topology"] = all(np.array_equal(p, q)
                                   for p, q in zip(dig["plain"], dig["learned"]))
    res["workload"] = ("clustered_4m: 4,194,304 periodic seeds, 80 % in 64 wrapped Gaussian "
                       "clusters (sigma log-uniform in [1e-3, 1e-1]), 20 % uniform")
    return res


def host_to_host(L, x, modes, reps=3):
    """mv_build from host seeds to host result arrays (the PCIe-inclusive
    rate): wall-clock median of reps builds after one warm-up build, and the
    bytes each build moves over the bus."""
    import numpy as np
    from meshless_voro_amd import _native
    from meshless_voro_amd.voronoi import VoronoiError
    from meshless_voro_amd.autotune import tuned_options
    n = len(x)
    xc = np.ascontiguousarray(x, dtype=np.float64)
    a, w = np.zeros(3), np.ones(3)
    res = {}
    for name, mode in modes:
        print(f"bench: host to host, mode {name}", file=sys.stderr, flush=True)
        opt = tuned_options(n, True, mode).to_c()
        ts = []
        for r in range(reps + 1):
            h = C.c_void_p()
            t0 = time.perf_counter()
            st = L.mv_build(xc.ctypes.data, n, None, a.ctypes.data, w.ctypes.data, 3, 1,
                            C.byref(opt), C.byref(h))
            dt = time.perf_counter() - t0
            if st != 0:
                raise VoronoiError(st)
            if r:
                ts.append(dt)
            if r == reps:
                F, NL = L.mv_n_faces(h), L.mv_n_links(h)
            L.mv_free(h)
        ms = 1e3 * float(np.median(ts))
        # cells: loc, centroid (3 f64 each), volume, safety radius (f64), face
        # offset, count and index (u64), entries used (u32); faces: left (u64),
        # right (i64), area (f64), centroid, normal, shift (3 f64 each),
        # has_shift (u8); links (u64)
        down = n * (24 + 24 + 8 + 8 + 24 + 4) + F * (8 + 8 + 8 + 72 + 1) + NL * 8
        res[name] = {"ms_per_build": ms, "cells_per_s": n / (ms / 1e3),
                     "upload_bytes": 24 * n, "download_bytes": down}
    return res


def build_device(L, xt, n, mode, periodic=True, kind="uniform"):
    from meshless_voro_amd import _native
    from meshless_voro_amd.voronoi import BuildOptions, VoronoiError
    import numpy as np
    from meshless_voro_amd.autotune import tuned_options
    a = np.zeros(3)
    w = np.ones(3)
    # launch and layout parameters from the autotuner's table (defaults if none)
    opt = tuned_options(n, periodic, mode, kind=kind).to_c()
    h = C.c_void_p()
    st = L.mv_build_device(C.c_void_p(xt.data_ptr()), n, None, a.ctypes.data, w.ctypes.data,
                           3, int(periodic), C.byref(opt), C.byref(h))
    if st != 0:
        raise VoronoiError(st)
    s = _native.mv_stats()
    L.mv_get_stats(h, C.byref(s))
    return h, {k: getattr(s, k) for k, _ in _native.mv_stats._fields_}


def connectivity_digest(L, h):
    """Host copies of the connectivity arrays (outside the timed region)."""
    import numpy as np
    F = L.mv_n_faces(h)
    NL = L.mv_n_links(h)
    def arr(p, n, ct):
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).copy()
    return (arr(L.mv_face_left(h), F, C.c_uint64), arr(L.mv_face_right(h), F, C.c_int64),
            arr(L.mv_cell_face_connections(h), NL, C.c_uint64))


def cpu_threads():
    """Threads for the CPU baseline: every core of this process's affinity
    mask, within the host share the environment grants (OMP_NUM_THREADS, set
    to the CPU share of one GPU on a shared GPU host)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    return (min(aff, int(share)) if share.isdigit() and int(share) > 0 else aff), aff


def cpu_baseline(n_sample, reps=3):
    """The oracle (a C restatement of the reference's construction) on the
    workload's seeds (the first n_sample of them), on this host's cores: one
    warm-up build, then the median of `reps` timed builds."""
    import numpy as np
    from oracle import oracle as O
    from meshless_voro_amd import workloads as W
    cores, aff = cpu_threads()
    x = W.uniform_2m()[:n_sample] if n_sample <= 2097152 else W.uniform(n_sample, 2)
    O.build(x, [0, 0, 0], [1, 1, 1], 3, True, nthreads=cores)  # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.build(x, [0, 0, 0], [1, 1, 1], 3, True, nthreads=cores)
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": n_sample / dt, "unit": "cells/s", "cores": cores, "kind": "port",
            "host_cores": os.cpu_count(), "affinity_cores": aff,
            "runs_s": [round(t, 3) for t in ts],
            "sample": f"{n_sample} i.i.d. uniform seeds (SplitMix64 stream 2"
                      f"{', the whole headline set' if n_sample == 2097152 else ''}), periodic "
                      f"unit cube, full build incl. grid/finalize; median of {reps} after a "
                      f"warm-up build: {dt:.2f} s on {cores} threads ({aff} cores in the "
                      f"affinity mask, {os.cpu_count()} on the host); CPU: {model}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU build would shard the seeds by slabs with a halo; not built)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from meshless_voro_amd import _native, workloads as W
    from meshless_voro_amd.voronoi import MODE_LEARNED, MODE_PLAIN

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    L = _native.lib()
    n = args.n
    x = W.uniform_2m() if n == 2097152 else W.uniform(n, 2)
    xt = torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0")
    torch.cuda.synchronize()

    modes = {}
    learned_note = None
    if args.modes in ("auto", "both", "plain"):
        modes["plain"] = MODE_PLAIN
    if args.modes in ("auto", "both", "learned"):
        try:
            from meshless_voro_amd import learned
            learned.default_model()
            modes["learned"] = MODE_LEARNED
        except Exception as e:  # no checkpoint: report the plain mode alone
            if args.modes == "learned":
                raise
            learned_note = f"learned mode unavailable: {e}"

    results = {}
    digests = {}
    for name, mode in modes.items():
        print(f"bench: mode {name}", file=sys.stderr, flush=True)
        for _ in range(args.warmup):
            h, _ = build_device(L, xt, n, mode)
            L.mv_free(h)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        stats = []
        t0 = time.perf_counter()
        e0.record()
        for _ in range(args.steps):
            h, st = build_device(L, xt, n, mode)
            stats.append(st)
            L.mv_free(h)
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1) / args.steps
        def tot(k):
            return float(np.sum([s[k] for s in stats]))
        results[name] = {"ms_per_step": ms, "wall_ms_per_step": 1e3 * wall / args.steps,
                         "cells_per_s": n / (ms / 1e3), "stats": stats[-1],
                         "ms_clip": float(np.mean([s["ms_clip"] for s in stats])),
                         "ms_knn": float(np.mean([s["ms_knn"] for s in stats])),
                         "clip_wave": {"launches": tot("clip_wave_launches"),
                                       "cells": tot("clip_wave_cells"),
                                       "ms": tot("ms_clip_wave")},
                         "knn": {"launches": tot("knn_launches"), "queries": tot("knn_queries"),
                                 "entries": tot("knn_entries"), "ms": tot("ms_knn_kernels")}}
        h, _ = build_device(L, xt, n, mode)
        digests[name] = connectivity_digest(L, h)
        L.mv_free(h)

    same = None
    if "learned" in digests and "plain" in digests:
        a, b = digests["plain"], digests["learned"]
        same = all(np.array_equal(p, q) for p, q in zip(a, b))
    best = max(results, key=lambda k: results[k]["cells_per_s"])
    R = results[best]

    # roofline of the dominant kernel of the best mode, per launch
    cw, kn = R["clip_wave"], R["knn"]
    if cw["ms"] >= kn["ms"] and cw["launches"] > 0:
        avg_ms = cw["ms"] / cw["launches"]
        units = cw["cells"] / cw["launches"]
        achieved = FLOPS_PER_CELL * units / (avg_ms / 1e3) / 1e12
        roof = {"bound": "valu", "pipe": "fp64 vector ALU (78.6 TFLOP/s counts an FMA as two "
                                          "flops; the kernel is compiled without contraction, "
                                          "so its op rate peaks at 39.3 Top/s)",
                "kernel": "k_clip_wave", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "peak_unfused": FP64_PEAK_TFLOPS / 2,
                "frac_unfused": achieved / (FP64_PEAK_TFLOPS / 2),
                "traffic": traffic_from_profile("k_clip_wave"),
                "per_unit": f"{FLOPS_PER_CELL:.0f} FP64 flops per completed cell",
                "units_per_launch": units, "avg_launch_ms": avg_ms,
                "launches": cw["launches"]}
    else:
        avg_ms = kn["ms"] / kn["launches"]
        q = kn["queries"] / kn["launches"]
        ent = kn["entries"] / kn["launches"]
        bytes_ = 32 * q + 8 * ent + 12 * q
        achieved = bytes_ / (avg_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "kernel": "k_knn", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic_from_profile("k_knn"),
                "per_unit": "32 B query position + 8 B per list entry + 12 B count/certificate",
                "units_per_launch": q, "avg_launch_ms": avg_ms, "launches": kn["launches"]}

    out = {
        "metric": METRIC, "value": R["cells_per_s"], "unit": "cells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": R["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "uniform_2m: 2,097,152 i.i.d. uniform seeds (SplitMix64 "
                               "stream 2), periodic unit cube, FP64, full build",
                   "n": n, "mode": best},
        "modes": {k: {"cells_per_s": v["cells_per_s"], "ms_per_step": v["ms_per_step"],
                      "ms_clip": v["ms_clip"], "ms_knn": v["ms_knn"],
                      "passes": v["stats"]["passes"], "cells_redone": v["stats"]["cells_redone"]}
                  for k, v in results.items()},
        "speedup_vs_plain": (R["cells_per_s"] / results["plain"]["cells_per_s"]
                             if "plain" in results else None),
        "same_topology": same,
        "roofline": roof,
    }
    if learned_note:
        out["note"] = learned_note
    if args.batched:
        bm = MODE_LEARNED if "learned" in modes else MODE_PLAIN
        print("bench: many-problem builds (uniform half)", file=sys.stderr, flush=True)
        out["batched_c5_uniform"] = batched_c5(L, bm, reps=3, only_uniform=True)
        print("bench: many-problem builds (all 512)", file=sys.stderr, flush=True)
        out["batched_c5"] = batched_c5(L, bm, reps=1)
        out["batched_c5"]["mode"] = out["batched_c5_uniform"]["mode"] = (
            "learned" if bm == MODE_LEARNED else "plain")
    if args.clustered:
        out["clustered_4m"] = clustered_4m(L, list(modes.items()))
    if args.host:
        out["host_to_host"] = host_to_host(L, x, list(modes.items()))
    if args.calibrate_copy:
        calibrate_copy()
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
