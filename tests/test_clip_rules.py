"""The wave clip kernel replaces two sequential steps of the reference's
clip_by_plane with rules that all lanes evaluate at once
(meshless_voro_amd/csrc/mv_clipw.hpp).  These CPU tests check each rule,
restated here in the kernel's terms, against the sequential step it stands
for:

  - the swap-with-last scan (src/voronoi/convex_cell.rs:368-390): the
    closed-form destination of every vertex slot equals the scan's result
    for every clip mask of up to 13 vertices and random masks up to 64;
  - compute_boundary (src/voronoi/convex_cell.rs:421-441 with
    SimpleCycle::try_extend, src/simple_cycle.rs): whenever the kernel's
    disk test accepts a set of removed dual triangles, the oracle's
    step-by-step SimpleCycle ends with the same cycle from the same start.
"""
import itertools
import random

import numpy as np
import pytest

from oracle import oracle as O


def reference_swap_scan(nv, clipped):
    """convex_cell.rs:368-390: scan from the front, swap each clipped vertex
    with the last unscanned one."""
    v = list(range(nv))
    i, num_v = 0, nv
    while i < num_v:
        if clipped[v[i]]:
            num_v -= 1
            v[i], v[num_v] = v[num_v], v[i]
        else:
            i += 1
    return v, num_v


def kernel_swap_rule(nv, clipped):
    """Destination slot of every vertex as k_clip_wave computes it: holes
    h_j (clipped, below K) and b_j (kept, from K up, ranked from the top)."""
    M = sum(1 << p for p in range(nv) if clipped[p])
    kept = ((1 << nv) - 1) & ~M
    K = bin(kept).count("1")
    low = (1 << K) - 1
    holes, backk = M & low, kept & ~low
    tab_h, tab_b = {}, {0: nv}
    for p in range(nv):
        if holes >> p & 1:
            tab_h[bin(holes & ((1 << p) - 1)).count("1") + 1] = p
        if backk >> p & 1:
            tab_b[bin(backk >> p).count("1")] = p
    dst = list(range(nv))
    for p in range(nv):
        if backk >> p & 1:
            dst[p] = tab_h[bin(backk >> p).count("1")]
        elif holes >> p & 1:
            dst[p] = tab_b[bin(holes & ((1 << p) - 1)).count("1")] - 1
        elif M >> p & 1:
            lowest_b = (backk & -backk).bit_length() - 1 if backk else nv
            dst[p] = lowest_b - 1 if p == K else p - 1
    out = [None] * nv
    for p in range(nv):
        assert out[dst[p]] is None, "not a permutation"
        out[dst[p]] = p
    return out, K


def test_swap_rule_all_small_masks():
    n = 0
    for nv in range(1, 14):
        for bits in itertools.product([False, True], repeat=nv):
            if any(bits):
                assert kernel_swap_rule(nv, bits) == reference_swap_scan(nv, bits)
                n += 1
    assert n == sum(2 ** k - 1 for k in range(1, 14))


def test_swap_rule_random_masks():
    rng = random.Random(11)
    for _ in range(20000):
        nv = rng.randint(1, 64)
        pr = rng.random()
        bits = [rng.random() < pr for _ in range(nv)]
        if any(bits):
            assert kernel_swap_rule(nv, bits) == reference_swap_scan(nv, bits)


def kernel_boundary_rule(tris):
    """The lanes' boundary: returns the cycle from the first plane of the
    first triangle, or None when the kernel falls back to the walk.  Plane x
    holds its out-edges (emat) and in-edges (rmat); its boundary out-edges
    are those whose reverse is absent."""
    emat, rmat = [0] * 64, [0] * 64
    for a, b, c in tris:
        for x, y in ((a, b), (b, c), (c, a)):
            emat[x] |= 1 << y
            rmat[y] |= 1 << x
    bo = [emat[x] & ~rmat[x] for x in range(64)]
    if any(bin(m).count("1") > 1 for m in bo):
        return None
    succ = [(m.bit_length() - 1) if m else None for m in bo]
    n_edge = sum(1 for m in bo if m)
    n_node = sum(1 for m in emat if m)
    n_dir = sum(bin(m).count("1") for m in emat)
    if n_dir != 3 * len(tris) or 2 * n_node != len(tris) + n_edge + 2:
        return None
    a0 = tris[0][0]
    cur, seq = a0, []
    for r in range(n_edge):
        nx = succ[cur]
        if nx is None or (nx == a0) != (r == n_edge - 1):
            return None
        seq.append(cur)
        cur = nx
    return seq


def sphere_triangulation(npts, rng):
    """Outward-oriented triangles of the convex hull of random points on the
    unit sphere: a closed, consistently oriented triangulated sphere, like
    the dual triangles of a convex cell's vertices."""
    from scipy.spatial import ConvexHull
    p = rng.normal(size=(npts, 3))
    p /= np.linalg.norm(p, axis=1)[:, None]
    inside = p.mean(axis=0)
    tri = []
    for s in ConvexHull(p).simplices:
        a, b, c = (int(i) for i in s)
        if np.dot(np.cross(p[b] - p[a], p[c] - p[a]), p[a] - inside) < 0:
            b, c = c, b
        tri.append((a, b, c))
    return p, tri


@pytest.mark.parametrize("orient", [1, -1])
def test_boundary_rule_matches_simple_cycle(orient):
    """Caps (triangles whose centre lies beyond a random plane) and grown
    regions of random triangulated spheres, in random order: where the rule
    accepts, the oracle's SimpleCycle succeeds with the same cycle."""
    rng = np.random.default_rng(5 + orient)
    accepted = agree_fail = 0
    for trial in range(400):
        p, tri = sphere_triangulation(int(rng.integers(6, 40)), rng)
        if orient < 0:
            tri = [(a, c, b) for a, b, c in tri]
        cen = np.array([(p[a] + p[b] + p[c]) / 3 for a, b, c in tri])
        if trial % 2 == 0:
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            h = rng.uniform(-0.5, 0.9)
            sel = [t for t, c in zip(tri, cen) if c @ d > h]
        else:
            # a region grown from one triangle through random shared edges:
            # a disk, or not (holes, pinched planes)
            sel = [tri[int(rng.integers(len(tri)))]]
            for _ in range(int(rng.integers(1, len(tri)))):
                t = sel[int(rng.integers(len(sel)))]
                nb = [u for u in tri if u not in sel and len(set(u) & set(t)) == 2]
                if nb:
                    sel.append(nb[int(rng.integers(len(nb)))])
        if not sel:
            continue
        sel = [sel[i] for i in rng.permutation(len(sel))]
        fast = kernel_boundary_rule(sel)
        seq = O.simple_cycle(sel, 64)
        if fast is not None:
            accepted += 1
            assert seq == fast
        elif seq is None:
            agree_fail += 1
    assert accepted > 100
