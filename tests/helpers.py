"""Shared checks for the parity tests."""
from __future__ import annotations

import math

import numpy as np

CONNECTIVITY = ("face_left", "face_right", "face_has_shift", "links", "cell_face_offset",
                "cell_face_count", "cell_idx")
INTEGRALS = ("cell_volume", "cell_centroid", "cell_safety_radius", "face_area",
             "face_centroid", "face_normal", "face_shift", "cell_loc")

# tolerance for FP64 integrals: relative 1e-12, scaled for near-zero values
# by the cell size (h^3 volumes, h^2 areas)
REL_TOL = 1e-12


def assert_connectivity_equal(a, b):
    for f in CONNECTIVITY:
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape, (f, x.shape, y.shape)
        assert np.array_equal(x, y), f


def assert_integrals_close(a, b, rel=REL_TOL):
    """|x - y| <= rel * max(|y|, scale) per field; scale = the typical cell
    size to the field's power, so that sub-ulp-of-the-cell noise on a
    near-zero area does not count as a relative error."""
    n = max(1, a.cell_volume.shape[0])
    vol_box = float(np.sum(np.abs(b.cell_volume))) or 1.0
    h = (vol_box / n) ** (1.0 / 3.0)
    scales = {"cell_volume": h ** 3, "face_area": h ** 2, "cell_safety_radius": h,
              "cell_centroid": h, "face_centroid": h, "cell_loc": h, "face_normal": 1.0,
              "face_shift": h}
    for f in INTEGRALS:
        x, y = getattr(a, f), getattr(b, f)
        assert x.shape == y.shape, f
        if x.size == 0:
            continue
        lim = rel * np.maximum(np.abs(y), scales[f])
        bad = np.abs(x - y) > lim
        assert not bad.any(), (f, int(bad.sum()), float(np.max(np.abs(x - y))))


def consistency_check(v, dim=3):
    """Voronoi::consistency_check (src/voronoi.rs:364-408): isoperimetric
    bound per active cell, and total volume = box volume when all cells are
    active (3D)."""
    total = 0.0
    all_active = True
    for i in range(v.cell_volume.shape[0]):
        vol = v.cell_volume[i]
        if vol == 0.0:
            all_active = False
            continue
        total += vol
        o, c = int(v.cell_face_offset[i]), int(v.cell_face_count[i])
        area = float(np.sum(v.face_area[v.links[o:o + c]]))
        if dim == 3:
            r = (0.25 * 3.0 / math.pi * vol) ** (1.0 / 3.0)
            assert area > 4.0 * math.pi * r * r
        elif dim == 2:
            r = math.sqrt(vol / math.pi)
            assert area > 2.0 * math.pi * r
        else:
            assert abs(area - 2.0) < 1e-9
    if all_active and dim == 3:
        w = v.width() if callable(v.width) else v.width
        box = w[0] * w[1] * w[2]
        assert abs(total - box) <= box * 1e-11 + 4 * np.spacing(box)
    return total


def closure_residual(v):
    """max over cells of |sum_f A_f n_f| / sum_f A_f with n_f oriented out of
    the cell; a closed polyhedron gives 0."""
    worst = 0.0
    for i in range(v.cell_volume.shape[0]):
        if v.cell_volume[i] == 0.0:
            continue
        o, c = int(v.cell_face_offset[i]), int(v.cell_face_count[i])
        fs = v.links[o:o + c].astype(np.int64)
        sgn = np.where(v.face_left[fs] == i, -1.0, 1.0)  # normal points to left
        s = np.sum((v.face_area[fs] * sgn)[:, None] * v.face_normal[fs], 0)
        worst = max(worst, float(np.linalg.norm(s) / np.sum(v.face_area[fs])))
    return worst


def degenerate_fixture(r):
    """The point set of the reference's degenerate_test /
    degenerate_test_simplified (src/voronoi.rs:1157-1184): perturbed_grid
    with pert 0 (src/voronoi.rs:840-859), i.e. per component
    (i + 0.5) * width / count + anchor, with one seed moved.  Returns
    (positions, anchor, width).  The bits are checked against that formula
    evaluated in Python floats, in the reference's order of operations."""
    from meshless_voro_amd import workloads as W
    count = int(r["grid"])
    anchor = np.array(r["anchor"], np.float64)
    width = np.array(r["width"], np.float64)
    x = W.perturbed_grid(count, 0.0, anchor=anchor, width=width)
    for n in range(count ** 3):
        ijk = (n // count ** 2, (n % count ** 2) // count, n % count)
        for c in range(3):
            want = (float(ijk[c]) + 0.5) * float(width[c]) / float(count) + float(anchor[c])
            want = min(max(want, float(anchor[c])), float(anchor[c]) + float(width[c]))
            assert x[n, c] == want, (n, c, x[n, c], want)
    x[r["moved_index"]] = r["moved_to"]
    return x, anchor, width
