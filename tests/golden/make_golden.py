"""Writes the golden vectors under tests/golden/.

Nothing here runs the reference: expected values come from mathematics
(Python's exact integers) and IEEE float64 arithmetic in numpy, and the
point sets of the reference's own tests are copied as data.

  insphere_vectors.json  5-tuples of 52-bit integer points and the sign of
                         det[b-a, c-a, d-a, v-a ; squared norms] computed
                         with Python ints by the Leibniz formula (the
                         quantity of in_sphere_test_exact, src/geometry.rs:
                         271-314): random, exactly co-spherical (snapped box
                         corners) and near-extreme coordinates.
  iloc_vectors.json      SimulationBoundary::iloc (src/voronoi/boundary.rs:
                         49-64) evaluated with numpy float64 (no FMA):
                         1 + (x - A) * iw, mantissa bits.
  reference_points.json  point sets of the reference tests
                         test_non_perturbed_z (src/voronoi.rs:1248-1295),
                         degenerate_test (:1157-1165) and
                         degenerate_test_simplified (:1167-1184).
Run: python tests/golden/make_golden.py
"""
import itertools
import json
import os
import random

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
M52 = (1 << 52) - 1


def det4_leibniz(m):
    tot = 0
    for p in itertools.permutations(range(4)):
        inv = sum(1 for i in range(4) for j in range(i + 1, 4) if p[i] > p[j])
        term = 1
        for r in range(4):
            term *= m[r][p[r]]
        tot += -term if inv % 2 else term
    return tot


def insphere_sign(a, b, c, d, v):
    cols = []
    for pnt in (b, c, d, v):
        diff = [pnt[i] - a[i] for i in range(3)]
        cols.append(diff + [sum(x * x for x in diff)])
    m = [[cols[j][i] for j in range(4)] for i in range(4)]
    det = det4_leibniz(m)
    return (det > 0) - (det < 0)


def gen_insphere(rng):
    out = []
    # random points anywhere in [0, 2^52)
    for _ in range(400):
        pts = [[rng.randrange(0, 1 << 52) for _ in range(3)] for _ in range(5)]
        out.append(pts)
    # exactly co-spherical: corners of an axis-aligned integer box
    for _ in range(300):
        lo = [rng.randrange(0, 1 << 51) for _ in range(3)]
        ext = [rng.randrange(1, 1 << 50) for _ in range(3)]
        corners = [[lo[i] + (ext[i] if (k >> i) & 1 else 0) for i in range(3)] for k in range(8)]
        pick = rng.sample(corners, 5)
        out.append(pick)
    # perturbed co-spherical (one coordinate moved by +-1)
    for _ in range(300):
        lo = [rng.randrange(1, 1 << 51) for _ in range(3)]
        ext = [rng.randrange(1, 1 << 50) for _ in range(3)]
        corners = [[lo[i] + (ext[i] if (k >> i) & 1 else 0) for i in range(3)] for k in range(8)]
        pick = [list(p) for p in rng.sample(corners, 5)]
        pick[4][rng.randrange(3)] += rng.choice((-1, 1))
        out.append(pick)
    # near-extreme magnitudes
    for _ in range(200):
        pts = [[rng.choice((rng.randrange(0, 16), M52 - rng.randrange(0, 16))) for _ in range(3)]
               for _ in range(5)]
        out.append(pts)
    # small local clusters (typical of neighbouring cells)
    for _ in range(300):
        base = [rng.randrange(1 << 40, 1 << 51) for _ in range(3)]
        pts = [[base[i] + rng.randrange(-(1 << 20), 1 << 20) for i in range(3)] for _ in range(5)]
        out.append(pts)
    return [{"pts": p, "sign": insphere_sign(*p)} for p in out]


def iloc_np(anchor, width, periodic, dim, loc):
    a = np.array(anchor, np.float64)
    w = np.array(width, np.float64)
    if periodic:
        a[0] -= w[0]; w[0] *= 3.0
        if dim >= 2:
            a[1] -= w[1]; w[1] *= 3.0
        if dim == 3:
            a[2] -= w[2]; w[2] *= 3.0
    A = a - w
    iw = 1.0 / (3.0 * w)
    t = 1.0 + (np.array(loc, np.float64) - A) * iw
    return [int(x) & M52 for x in t.view(np.uint64)]


def gen_iloc(rng):
    out = []
    for _ in range(500):
        anchor = [rng.uniform(-100, 100) for _ in range(3)]
        width = [rng.uniform(0.01, 50) for _ in range(3)]
        periodic = rng.random() < 0.5
        loc = [anchor[i] + rng.random() * width[i] for i in range(3)]
        out.append({"anchor": anchor, "width": width, "periodic": periodic, "dim": 3,
                    "loc": loc, "iloc": iloc_np(anchor, width, periodic, 3, loc)})
    return out


REFERENCE_POINTS = {
    "test_non_perturbed_z": {
        "anchor": [-14.17150624593099, 16.202089309692383, 26.540990829467773],
        "width": [32.670213063557945, 4.626067479451496, 0.2923425038655587],
        "points": [
            [17.159128189086914, 17.747779846191406, 26.790985743204754],
            [16.933035532633465, 17.705819447835285, 26.666666666666668],
            [17.00386683146159, 17.983611424763996, 26.833333333333332],
            [16.84653155008952, 18.130948384602863, 26.833333333333332],
            [16.59058952331543, 18.026987075805664, 26.666666666666668],
            [16.594066619873047, 18.36242739359538, 26.833333333333332],
            [16.32268778483073, 18.27001889546712, 26.666666666666668],
            [16.346779505411785, 18.498706817626953, 26.790985743204754],
            [16.202088673909504, 18.457736333211262, 26.70674769083659],
            [16.058500289916992, 18.417078018188477, 26.623419443766277],
            [15.915996233622232, 18.376726786295574, 26.540990829467773],
            [17.12963612874349, 17.60032081604004, 26.70674769083659],
            [17.100369135538738, 17.4539852142334, 26.623419443766277],
            [17.071322758992512, 17.308754603068035, 26.540990829467773],
            [-14.153460820515951, 19.766197840372723, 26.540990829467773],
            [-14.159430185953775, 19.91418393452962, 26.623419443766277],
            [-14.165445327758789, 20.063296635945637, 26.70674769083659],
            [-14.17150624593099, 20.21355374654134, 26.790985743204754],
            [-13.954760869344076, 20.136741638183594, 26.666666666666668],
            [-13.981264114379883, 20.42219416300456, 26.833333333333332],
            [-13.655904134114584, 20.340506235758465, 26.666666666666668],
            [-13.697244962056478, 20.613646189371746, 26.833333333333332],
            [-13.51724942525228, 20.73223940531413, 26.833333333333332],
            [-13.263668060302734, 20.598515192667644, 26.666666666666668],
            [-13.251688321431478, 20.82815678914388, 26.790985743204754],
            [-13.115188280741373, 20.765055974324543, 26.70674769083659],
            [-12.979728062947592, 20.702435811360676, 26.623419443766277],
            [-12.845291455586752, 20.640288670857746, 26.540990829467773],
            [18.498706817626953, 16.34678014119466, 26.790985743204754],
            [18.270018259684246, 16.322688420613606, 26.666666666666668],
            [18.3624267578125, 16.594066619873047, 26.833333333333332],
            [18.21713638305664, 16.75329335530599, 26.833333333333332],
            [17.953826268513996, 16.66973304748535, 26.666666666666668],
            [17.983611424763996, 17.00386683146159, 26.833333333333332],
            [17.705819447835285, 16.933035532633465, 26.666666666666668],
            [17.747779846191406, 17.159128189086914, 26.790985743204754],
            [17.60032081604004, 17.12963612874349, 26.70674769083659],
            [17.4539852142334, 17.100369135538738, 26.623419443766277],
            [17.308754603068035, 17.071322758992512, 26.540990829467773],
            [18.457736333211262, 16.202089309692383, 26.70674769083659],
        ],
    },
    "degenerate_test": {
        "anchor": [0.0, 0.0, 0.0], "width": [2e15, 2e15, 2e15], "grid": 10, "moved_index": 42,
        "moved_to": [1e14 * 1.00007490802, 1e14 * 9.00019014286, 1e14 * 5.00014639879],
    },
    "degenerate_test_simplified": {
        "anchor": [0.0, 6e14, 2e14], "width": [6e14, 6e14, 6e14], "grid": 3, "moved_index": 4,
        "moved_to": [1e14 * 1.00007490802, 1e14 * 9.00019014286, 1e14 * 5.00014639879],
        "mask_true": [3],
    },
}


def main():
    rng = random.Random(20240521)
    with open(os.path.join(HERE, "insphere_vectors.json"), "w") as f:
        json.dump(gen_insphere(rng), f)
    with open(os.path.join(HERE, "iloc_vectors.json"), "w") as f:
        json.dump(gen_iloc(rng), f)
    with open(os.path.join(HERE, "reference_points.json"), "w") as f:
        json.dump(REFERENCE_POINTS, f, indent=1)


if __name__ == "__main__":
    main()
