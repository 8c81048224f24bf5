"""Cells beyond the large-cell kernel's capacity (255 planes, 1024 vertices)
run on the wide kernel (device-memory arrays, 21-bit plane indices,
capacity doubled on demand), where the reference grows its arrays without
bound (src/voronoi/convex_cell.rs:394-418).  A seed surrounded by hundreds
of seeds on a sphere has one face per sphere seed; clusters in empty space
have rim cells of hundreds of vertices, which lowered capacities of the
large-cell kernel (a build option) send to the wide kernel.  Every case
equals the oracle."""
import numpy as np
import pytest

from meshless_voro_amd import BuildOptions, Voronoi, VoronoiError
from oracle import oracle as O
from tests.helpers import assert_connectivity_equal, assert_integrals_close

pytestmark = pytest.mark.gpu
ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]


def sphere_cell(m, seed):
    """One seed at the centre, m seeds near a sphere of radius 0.2 around
    it (radii jittered by 1e-3, so no five seeds are co-spherical), and a
    uniform background beyond radius 0.35."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    shell = 0.5 + d * (0.2 * (1.0 + 1e-3 * rng.random((m, 1))))
    bg = rng.random((4000, 3))
    bg = bg[np.linalg.norm(bg - 0.5, axis=1) > 0.35]
    return np.concatenate([[[0.5, 0.5, 0.5]], shell, bg])


def both(x, periodic, options=None):
    print(f"case wide n={len(x)} periodic={periodic}", flush=True)
    g = Voronoi.build(x, ZERO, ONE, 3, periodic, options=options)
    o = O.build(x, ZERO, ONE, 3, periodic)
    assert_connectivity_equal(g, o)
    assert_integrals_close(g, o)
    assert np.array_equal(g.cell_n_popped, o.cell_n_popped)
    return g, o


@pytest.mark.parametrize("m", [600, 1500])
def test_cell_over_the_large_cell_capacity(m):
    """m = 600: ~600 planes, ~1200 vertices (over 255 and 1024); m = 1500:
    over the wide kernel's first capacity too (1024 planes), which then
    doubles."""
    x = sphere_cell(m, 5 + m)
    g, o = both(x, False)
    assert int(g.cell_face_count[0]) > 255
    assert g.stats["wide_cells"] >= 1


@pytest.mark.parametrize("periodic", [False, True])
def test_lowered_capacity_forces_the_wide_kernel(periodic):
    rng = np.random.default_rng(79)
    n = 6000
    c = rng.random((8, 3))
    x = np.mod(c[rng.integers(0, 8, n)] + 0.03 * rng.normal(size=(n, 3)), 1.0)
    x[x >= 1.0] = 0.0
    g, o = both(x, periodic, BuildOptions(coop_max_planes=40, coop_max_vertices=72))
    assert g.stats["big_cells"] > 0 and g.stats["wide_cells"] > 0


def test_kept_cells_over_capacity_report_it():
    """Kept-cell builds (VoronoiIntegrator) store 8-bit plane indices: a cell
    over the large-cell kernel's capacity returns MV_E_CAPACITY there."""
    from meshless_voro_amd.integrator import VoronoiIntegrator
    x = sphere_cell(600, 9)
    with pytest.raises(VoronoiError) as e:
        VoronoiIntegrator.build(x, None, ZERO, ONE, 3, False)
    assert e.value.status == 9
