"""Loading and building libmvoro.so (the gfx950 HIP library behind the C ABI
of include/mvoro.h).

There is no CPU fallback: when the library is missing or no GPU is present,
the entry points raise.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

_PKG = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_PKG)
LIB_PATH = os.path.join(_PKG, "libmvoro.so")
SOURCES = [os.path.join(_PKG, "csrc", f) for f in (
    "mvoro.hip", "mv_common.hpp", "mv_geom.hpp", "mv_grid.hpp", "mv_clip.hpp",
    "mv_clipw.hpp", "mv_faces.hpp", "mv_finalize.hpp", "mv_learn.hpp")] + [
    os.path.join(_ROOT, "include", f) for f in ("mvoro.h", "mvoro_integrals.hpp")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # the reference never fuses multiply-add; clip decisions must see the
    # same bits
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose: bool = False) -> str:
    """Compile libmvoro.so in-tree with hipcc for gfx950."""
    cmd = ["hipcc", *HIPCC_FLAGS, "-I", os.path.join(_ROOT, "include"),
           "-I", os.path.join(_PKG, "csrc"),
           "-o", LIB_PATH + ".tmp", os.path.join(_PKG, "csrc", "mvoro.hip")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    if verbose:
        print(" ".join(cmd))
    return LIB_PATH


_lib = None


class mv_options(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("k0", C.c_uint32), ("k_max", C.c_uint32),
                ("lanes", C.c_uint32), ("grid_occupancy16", C.c_uint32),
                ("flags", C.c_uint32), ("model", C.c_void_p),
                ("clip_blocks_per_cu", C.c_uint32),
                ("coop_max_planes", C.c_uint32), ("coop_max_vertices", C.c_uint32),
                ("cells_per_wave", C.c_uint32), ("first_list_cap", C.c_uint32),
                ("knn_min_tile", C.c_uint32), ("pipeline", C.c_uint32)]


MV_MODEL_MAGIC = 0x4B4C564D
MV_FLAG_EXACT_ORDER = 1
MV_FLAG_KEEP_CELLS = 2
MV_FLAG_NO_LENGTH_ORDER = 4


class mv_cell_geometry(C.Structure):
    _fields_ = [("pl_a", C.c_void_p), ("pl_b", C.c_void_p), ("vx", C.c_void_p),
                ("pbase", C.c_void_p), ("vbase", C.c_void_p), ("np", C.c_void_p),
                ("nv", C.c_void_p), ("cell_loc", C.c_void_p), ("width", C.c_double * 3),
                ("dim", C.c_int32), ("idx_base", C.c_uint64), ("fbase", C.c_void_p),
                ("nf", C.c_void_p), ("face_plane", C.c_void_p), ("face_nvert", C.c_void_p),
                ("face_voff", C.c_void_p), ("face_vtx", C.c_void_p)]


class mv_model(C.Structure):
    _fields_ = [("magic", C.c_uint32), ("version", C.c_uint32),
                ("mu", C.c_float * 16), ("sigma", C.c_float * 16),
                ("w1", C.c_uint16 * (16 * 64)), ("w2", C.c_uint16 * (64 * 64)),
                ("w3", C.c_uint16 * (64 * 2)), ("b1", C.c_float * 64),
                ("b2", C.c_float * 64), ("b3", C.c_float * 2),
                ("tau", C.c_float), ("kscale", C.c_float), ("rscale", C.c_float)]


class mv_train_config(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("batch", C.c_uint32), ("seed", C.c_uint32),
                ("lr", C.c_float), ("tau", C.c_float)]


class mv_stats(C.Structure):
    _fields_ = [("passes", C.c_uint32), ("k_last", C.c_uint32),
                ("cells_redone", C.c_uint64), ("exact_tests", C.c_uint64),
                ("big_cells", C.c_uint64), ("ms_grid", C.c_double),
                ("ms_knn", C.c_double), ("ms_clip", C.c_double),
                ("ms_finalize", C.c_double), ("ms_total", C.c_double),
                ("clip_wave_launches", C.c_uint32), ("knn_launches", C.c_uint32),
                ("clip_wave_cells", C.c_uint64), ("knn_queries", C.c_uint64),
                ("knn_entries", C.c_uint64),
                ("ms_clip_wave", C.c_double), ("ms_knn_kernels", C.c_double),
                ("ms_model", C.c_double), ("wide_cells", C.c_uint64),
                ("pair_handovers", C.c_uint64), ("knn_big_items", C.c_uint64),
                ("knn_big_gathers", C.c_uint64), ("knn_big_entries", C.c_uint64),
                ("knn_gathers", C.c_uint64), ("knn_reads", C.c_uint64),
                ("dense_cells", C.c_uint64)]


class mv_device_view(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "cell_loc", "cell_centroid", "cell_volume", "cell_safety_radius",
        "cell_face_offset", "cell_face_count", "cell_idx", "face_left",
        "face_right", "face_area", "face_centroid", "face_normal",
        "face_has_shift", "face_shift", "cell_face_connections")]


class mv_noalloc_out(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "cell_loc", "cell_centroid", "cell_volume", "cell_safety_radius",
        "cell_face_offset", "cell_face_count", "cell_idx", "face_left",
        "face_right", "face_area", "face_centroid", "face_normal",
        "face_has_shift", "face_shift", "cell_face_connections")] + [
        (n, C.c_size_t) for n in ("face_cap", "link_cap", "n_faces", "n_links",
                                  "scratch_used")]


NOALLOC_SCRATCH_BYTES = 128 << 20
NOALLOC_MAX_SEEDS = 32768


# every symbol declared in include/mvoro.h
EXPORTS = [
    "mv_default_options", "mv_build", "mv_build_device", "mv_build_batch",
    "mv_build_batch_device", "mv_free", "mv_anchor",
    "mv_width", "mv_dimensionality", "mv_periodic", "mv_n_cells", "mv_n_faces",
    "mv_n_links", "mv_cell_loc", "mv_cell_centroid", "mv_cell_volume",
    "mv_cell_safety_radius", "mv_cell_face_offset", "mv_cell_face_count",
    "mv_cell_idx", "mv_face_left", "mv_face_right", "mv_face_area",
    "mv_face_centroid", "mv_face_normal", "mv_face_has_shift", "mv_face_shift",
    "mv_cell_face_connections", "mv_get_device_view", "mv_get_stats",
    "mv_cell_n_popped", "mv_status_string", "mv_n_active", "mv_cell_integrals",
    "mv_face_integrals", "mv_tets", "mv_get_cell_geometry", "mv_model_features", "mv_model_train",
    "mv_model_predict", "mv_model_loss_grad", "mv_in_sphere_exact", "mv_kept_cells",
    "mv_with_faces", "mv_discard_faces", "mv_face_vertices", "mv_voronoi_from_integrator",
    "mv_test_run_shares", "mv_build_device_noalloc", "mv_test_device_mallocs",
    "mv_graph_create", "mv_graph_launch", "mv_graph_free", "mv_graph_last_violation",
]


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(hipcc, gfx950); there is no CPU fallback")
    # torch's HIP runtime first: its libraries ask for "libamdhip64.so" and
    # would load a second runtime beside one this library brought in (under
    # its versioned name), and device pointers of one are unknown to the
    # other.  Loaded after torch, the library binds to torch's runtime.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    vp, sz, u32 = C.c_void_p, C.c_size_t, C.c_uint32
    L.mv_default_options.argtypes = [C.POINTER(mv_options)]
    L.mv_build.restype = C.c_int
    L.mv_build.argtypes = [vp, sz, vp, vp, vp, u32, u32, C.POINTER(mv_options), C.POINTER(vp)]
    L.mv_build_device.restype = C.c_int
    L.mv_build_device.argtypes = [vp, sz, vp, vp, vp, u32, u32, C.POINTER(mv_options), C.POINTER(vp)]
    L.mv_build_batch.restype = C.c_int
    L.mv_build_batch.argtypes = [vp, vp, u32, vp, vp, vp, vp, u32, C.POINTER(mv_options), vp]
    L.mv_build_batch_device.restype = C.c_int
    L.mv_build_batch_device.argtypes = [vp, vp, u32, vp, vp, vp, vp, u32,
                                        C.POINTER(mv_options), vp]
    L.mv_free.argtypes = [vp]
    L.mv_graph_create.restype = C.c_int
    L.mv_graph_create.argtypes = [vp, vp, u32, vp, vp, vp, vp, u32, C.POINTER(mv_options),
                                  C.POINTER(vp)]
    L.mv_graph_launch.restype = C.c_int
    L.mv_graph_launch.argtypes = [vp, vp, C.POINTER(u32)]
    L.mv_graph_free.argtypes = [vp]
    L.mv_graph_last_violation.restype = u32
    L.mv_graph_last_violation.argtypes = [vp]
    L.mv_anchor.argtypes = [vp, vp]
    L.mv_width.argtypes = [vp, vp]
    for f in ("mv_dimensionality", "mv_periodic"):
        getattr(L, f).restype = u32
        getattr(L, f).argtypes = [vp]
    for f in ("mv_n_cells", "mv_n_faces", "mv_n_links"):
        getattr(L, f).restype = sz
        getattr(L, f).argtypes = [vp]
    for f in EXPORTS:
        if f.startswith(("mv_cell_", "mv_face_")):
            getattr(L, f).restype = vp
            getattr(L, f).argtypes = [vp]
    L.mv_get_device_view.argtypes = [vp, C.POINTER(mv_device_view)]
    L.mv_get_stats.argtypes = [vp, C.POINTER(mv_stats)]
    L.mv_status_string.restype = C.c_char_p
    L.mv_status_string.argtypes = [C.c_int]
    L.mv_n_active.restype = sz
    L.mv_n_active.argtypes = [vp]
    L.mv_cell_integrals.restype = C.c_int
    L.mv_cell_integrals.argtypes = [vp, u32, vp]
    L.mv_face_integrals.restype = C.c_int
    L.mv_face_integrals.argtypes = [vp, u32, u32, C.POINTER(C.c_size_t), vp, vp, vp, vp, vp, vp]
    L.mv_tets.restype = C.c_int
    L.mv_tets.argtypes = [vp, C.POINTER(C.c_size_t), vp, vp, vp]
    L.mv_kept_cells.restype = C.c_int
    L.mv_kept_cells.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)] + [vp] * 9
    L.mv_with_faces.restype = C.c_int
    L.mv_with_faces.argtypes = [vp]
    L.mv_discard_faces.restype = C.c_int
    L.mv_discard_faces.argtypes = [vp]
    L.mv_face_vertices.restype = C.c_int
    L.mv_face_vertices.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)] + [vp] * 5
    L.mv_voronoi_from_integrator.restype = C.c_int
    L.mv_voronoi_from_integrator.argtypes = [vp, C.POINTER(vp)]
    L.mv_get_cell_geometry.restype = C.c_int
    L.mv_get_cell_geometry.argtypes = [vp, C.POINTER(mv_cell_geometry)]
    L.mv_model_features.restype = C.c_int
    L.mv_model_features.argtypes = [vp, sz, vp, vp, u32, u32, vp, vp]
    L.mv_model_train.restype = C.c_int
    L.mv_model_train.argtypes = [vp, vp, sz, C.POINTER(mv_train_config),
                                 C.POINTER(mv_model), C.POINTER(C.c_float)]
    L.mv_model_predict.restype = C.c_int
    L.mv_model_predict.argtypes = [C.POINTER(mv_model), vp, sz, vp]
    L.mv_model_loss_grad.restype = C.c_int
    L.mv_model_loss_grad.argtypes = [C.POINTER(mv_model), vp, vp, sz, C.c_float,
                                     C.POINTER(C.c_float), vp]
    L.mv_in_sphere_exact.restype = C.c_int
    L.mv_in_sphere_exact.argtypes = [vp, sz, vp]
    L.mv_build_device_noalloc.restype = C.c_int
    L.mv_build_device_noalloc.argtypes = [vp, sz, vp, vp, vp, u32, u32, C.POINTER(mv_options),
                                          C.POINTER(mv_noalloc_out)]
    L.mv_test_device_mallocs.restype = C.c_uint64
    L.mv_test_device_mallocs.argtypes = []
    L.mv_test_run_shares.restype = C.c_int
    L.mv_test_run_shares.argtypes = [C.c_int32, C.c_int32, vp]
    _lib = L
    return L
