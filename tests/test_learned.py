"""Learned neighbour-count proposal: the NumPy reference pinned by finite
differences (CPU), the checkpoint format (CPU), and on the GPU the MFMA
inference and training kernels against the reference, the feature kernel's
invariants, and learned builds against plain builds and the oracle (same
cells and faces; fewer redone cells)."""
import os

import numpy as np
import pytest

from meshless_voro_amd import BuildOptions, Voronoi, workloads as W
from meshless_voro_amd import learned
from meshless_voro_amd.voronoi import MODE_LEARNED, MODE_PLAIN
from oracle import oracle as O
from tests import mlp_reference as R
from tests.helpers import assert_connectivity_equal, assert_integrals_close

ZERO, ONE = [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]


# ------------------------------------------------------------------ CPU
def test_reference_gradient_matches_finite_differences():
    """The FP64 reference gradient agrees with central differences of the
    FP64 loss (so the GPU backward is checked against a pinned formula)."""
    Wt = R.random_weights(3)
    rng = np.random.default_rng(4)
    F = rng.normal(0, 1, (64, R.NF)).astype(np.float32) * Wt["sigma"] + Wt["mu"]
    Y0, _ = R.forward(Wt, F, quantize=False)
    T = Y0 + rng.normal(0, 0.5, Y0.shape)
    tau = 0.8
    _, g = R.loss_grad(Wt, F, T, tau, quantize=False)
    theta = R.flatten(Wt)
    idx = rng.choice(R.NPARAM, 40, replace=False)
    eps = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += eps
        tm[i] -= eps
        fp = R.pinball(R.forward(R.unflatten(Wt, tp), F, False)[0], T, tau)
        fm = R.pinball(R.forward(R.unflatten(Wt, tm), F, False)[0], T, tau)
        fd = (fp - fm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_bf16_rounding_reference():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.14159, 0.0], np.float32)
    y = R.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 1.5*2^-8 -> 1 + 2^-7
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == np.float32(1.0078125)
    assert y[3] == np.float32(1.0078125) and y[4] == np.float32(-3.140625) and y[5] == 0.0


def test_checkpoint_format():
    """The committed checkpoint: magic, version, size, finite weights."""
    import ctypes as C
    from meshless_voro_amd import _native
    if not os.path.exists(learned.CHECKPOINT):
        pytest.skip("no checkpoint committed")
    assert os.path.getsize(learned.CHECKPOINT) == C.sizeof(_native.mv_model)
    m = learned.Model.load()
    w = m.weights()
    for k, v in w.items():
        assert np.all(np.isfinite(v)), k
    assert 0.5 < m.tau < 1.0 and m.kscale > 0
    assert np.all(w["sigma"] > 0)


def test_from_weights_roundtrip(tmp_path):
    Wt = R.random_weights(5)
    m = learned.Model.from_weights(Wt, tau=0.9, kscale=1.25)
    p = tmp_path / "m.mvlk"
    m.save(str(p))
    m2 = learned.Model.load(str(p))
    w = m2.weights()
    for k in ("w1", "w2", "w3"):
        assert np.array_equal(w[k], R.bf16(Wt[k])), k
    for k in ("mu", "sigma", "b1", "b2", "b3"):
        assert np.array_equal(w[k], Wt[k]), k
    assert m2.tau == np.float32(0.9) and m2.kscale == 1.25


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_inference_matches_reference():
    """k_mlp_infer (fused MFMA 16x16x16 BF16, LDS transposes) against the
    NumPy network with the same BF16 roundings."""
    Wt = R.random_weights(11)
    m = learned.Model.from_weights(Wt)
    rng = np.random.default_rng(12)
    n = 1000  # not a multiple of 16: exercises the ragged last tile
    F = (rng.normal(0, 1, (n, R.NF)) * Wt["sigma"] + Wt["mu"]).astype(np.float32)
    Yg = m.predict(F)
    Yr, _ = R.forward(Wt, F, quantize=True)
    err = np.abs(Yg - Yr)
    scale = 1.0 + np.abs(Yr)
    assert np.quantile(err / scale, 0.99) < 1e-4, np.quantile(err / scale, 0.99)
    assert np.max(err / scale) < 2e-2, np.max(err / scale)


@pytest.mark.gpu
def test_gradient_check():
    """One pass of the training kernels (BF16 MFMA GEMMs, pinball loss,
    ReLU masks, column sums) against the reference gradient."""
    Wt = R.random_weights(21)
    m = learned.Model.from_weights(Wt, tau=0.8)
    rng = np.random.default_rng(22)
    n = 512
    F = (rng.normal(0, 1, (n, R.NF)) * Wt["sigma"] + Wt["mu"]).astype(np.float32)
    Yr, _ = R.forward(Wt, F, True)
    T = (Yr + rng.normal(0, 0.5, Yr.shape)).astype(np.float32)
    loss_g, g_g = m.loss_grad(F, T, tau=0.8)
    Wq = dict(Wt)
    for k in ("w1", "w2", "w3"):
        Wq[k] = R.bf16(Wt[k])
    loss_r, g_r = R.loss_grad(Wq, F, T, 0.8, quantize=True)
    assert abs(loss_g - loss_r) <= 1e-4 * (1 + abs(loss_r))
    for k, (a, b) in R.SLICES.items():
        d = np.linalg.norm(g_g[a:b] - g_r[a:b]) / max(np.linalg.norm(g_r[a:b]), 1e-30)
        assert d < 1.5e-2, (k, d)


@pytest.mark.gpu
def test_training_fits_a_quantile():
    """Training on a known target: the learned tau-quantile covers about tau
    of the rows and the loss falls well below that of a constant."""
    rng = np.random.default_rng(31)
    n = 32768
    F = rng.normal(0, 1, (n, R.NF)).astype(np.float32)
    t0 = 0.8 * F[:, 0] - 0.5 * np.abs(F[:, 1]) + 0.1 * rng.normal(0, 1, n)
    t1 = np.tanh(F[:, 2]) + 0.05 * rng.normal(0, 1, n)
    T = np.stack([t0, t1], 1).astype(np.float32)
    tau = 0.9
    m, loss = learned.train(F, T, steps=1500, batch=1024, seed=3, lr=3e-3, tau=tau)
    const = sum(R.pinball(np.full((n, 1), np.quantile(T[:, c], tau)), T[:, c:c + 1], tau)
                for c in range(2))
    assert loss < 0.35 * const, (loss, const)
    Y = m.predict(F)
    cover = np.mean(T <= Y, axis=0)
    assert np.all(np.abs(cover - tau) < 0.04), cover


@pytest.mark.gpu
def test_features_permutation_invariant():
    x = W.uniform(5000, 41)
    F, h = learned.features(x, ZERO, ONE, 3, True)
    p = np.random.default_rng(42).permutation(len(x))
    F2, h2 = learned.features(x[p], ZERO, ONE, 3, True)
    assert np.array_equal(F[p], F2) and np.array_equal(h[p], h2)
    assert np.all(np.isfinite(F)) and np.all(h > 0)


@pytest.mark.gpu
@pytest.mark.parametrize("periodic", [True, False])
def test_features_dense_clusters_find_the_nearest(periodic):
    """Clusters of 10,000 seeds with sigma 2e-3 put thousands of seeds in one
    grid cell (two clusters sit at the periodic z wrap).  A wave of 64 seeds
    in one such cell shares the features of its first seed (the
    representative, whose candidates the 64 lanes scan in parallel); every
    other seed gets its own.  So each seed's 8 distances are the 8 nearest
    found by a k-d tree over all seeds, either its own or those of a seed
    with the same feature row (at most 64 share one), most dense seeds share
    theirs.  (Which seeds share a wave follows the grid's order of a cell,
    z then generator index, so a permutation of the input may regroup seeds
    of equal z key; the same properties hold for it.)"""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(17 + periodic)
    centres = np.array([[0.5, 0.5, 0.5], [0.3, 0.7, 0.004], [0.8, 0.2, 0.997],
                        [0.1, 0.1, 0.5]])
    x = np.concatenate([rng.random((20000, 3))] +
                       [c + 2e-3 * rng.standard_normal((10000, 3)) for c in centres])
    x = x % 1.0 if periodic else np.clip(x, 0.0, 1.0 - 1e-12)
    d, _ = cKDTree(x, boxsize=1.0 if periodic else None).query(x, k=9)
    d = d[:, 1:]
    dense = d[:, 7] < 2e-3
    assert dense.sum() > 35000
    for p in (np.arange(len(x)), rng.permutation(len(x))):
        F, h = learned.features(x[p], ZERO, ONE, 3, periodic)
        dp, densep = d[p], dense[p]
        got = F[:, :8].astype(np.float64) * h[:, None].astype(np.float64)
        own = np.all(np.abs(got - dp) <= 2e-6 * dp, axis=1)
        rows = np.concatenate([F, h[:, None]], 1)
        _, inv, cnt = np.unique(rows, axis=0, return_inverse=True, return_counts=True)
        inv = inv.reshape(-1)
        group_ok = np.zeros(len(cnt), bool)
        np.logical_or.at(group_ok, inv, own)
        # (a seed outside the clusters may have nearer seeds beyond the 3^3
        # block of grid cells its features search, so only dense seeds are
        # checked against the k-d tree)
        assert np.all(group_ok[inv[densep]])
        assert np.all(cnt <= 64)
        assert np.all(own[densep & (cnt[inv] == 1)])
        assert np.mean(cnt[inv[densep]] > 1) > 0.5


@pytest.mark.gpu
def test_features_perfect_grid():
    """On a periodic perfect grid the 6 nearest are at the spacing and the
    next 2 at spacing * sqrt(2)."""
    c = 16
    x = W.grid(c)
    F, h = learned.features(x, ZERO, ONE, 3, True)
    s = 1.0 / c
    d = F[:, :8] * h[:, None]
    assert np.allclose(d[:, :6], s, rtol=1e-5)
    assert np.allclose(d[:, 6:8], s * np.sqrt(2), rtol=1e-5)
    assert np.all(F[:, 14] == 1.0)


def _quick_model(xs, periodic):
    Fs, Ts = [], []
    for x in xs:
        F, h = learned.features(x, ZERO, ONE, 3, periodic)
        v = Voronoi.build(x, ZERO, ONE, 3, periodic)
        Fs.append(F)
        Ts.append(learned.targets(v, h))
    m, _ = learned.train(np.concatenate(Fs), np.concatenate(Ts), steps=600, batch=1024,
                         seed=7, lr=3e-3, tau=0.95)
    return m


@pytest.mark.gpu
@pytest.mark.parametrize("periodic", [False, True])
def test_learned_build_same_as_oracle(periodic):
    """A learned build returns exactly the oracle's cells and faces."""
    m = _quick_model([W.uniform(40000, 50 + i) for i in range(2)], periodic)
    x = W.uniform(20000, 60)
    print(f"case learned n={len(x)} periodic={periodic}", flush=True)
    g = Voronoi.build(x, ZERO, ONE, 3, periodic, options=BuildOptions(mode=MODE_LEARNED, model=m))
    o = O.build(x, ZERO, ONE, 3, periodic)
    assert_connectivity_equal(g, o)
    assert_integrals_close(g, o)
    assert np.array_equal(g.cell_n_popped, o.cell_n_popped)


@pytest.mark.gpu
def test_learned_equals_plain_and_redoes_fewer_cells():
    """Held-out seeds: identical output to the plain mode (bit for bit), and
    far fewer cells redone."""
    m = _quick_model([W.uniform(100000, 70 + i) for i in range(2)], True)
    x = W.uniform(200000, 80)
    print(f"case learned-vs-plain n={len(x)}", flush=True)
    p = Voronoi.build(x, ZERO, ONE, 3, True, options=BuildOptions(mode=MODE_PLAIN))
    g = Voronoi.build(x, ZERO, ONE, 3, True, options=BuildOptions(mode=MODE_LEARNED, model=m))
    for a in ("face_left", "face_right", "links", "cell_face_offset", "cell_face_count",
              "cell_volume", "cell_centroid", "face_area", "face_centroid", "face_normal",
              "cell_safety_radius", "cell_n_popped"):
        assert np.array_equal(getattr(p, a), getattr(g, a)), a
    assert g.stats["cells_redone"] < 0.25 * p.stats["cells_redone"], (g.stats, p.stats)


@pytest.mark.gpu
def test_learned_clustered_and_partial():
    """Clustered seeds (cells needing thousands of neighbours) and a partial
    build through the learned mode match the plain mode."""
    x = W.clustered(60000, 91)
    m = _quick_model([W.clustered(60000, 92)], True)
    mask = W.partial_mask(len(x), 0.1, 93)
    print(f"case learned clustered n={len(x)}", flush=True)
    for msk in (None, mask):
        if msk is None:
            p = Voronoi.build(x, ZERO, ONE, 3, True)
            g = Voronoi.build(x, ZERO, ONE, 3, True,
                              options=BuildOptions(mode=MODE_LEARNED, model=m))
        else:
            p = Voronoi.build_partial(x, msk, ZERO, ONE, 3, True)
            g = Voronoi.build_partial(x, msk, ZERO, ONE, 3, True,
                                      options=BuildOptions(mode=MODE_LEARNED, model=m))
        for a in ("face_left", "face_right", "links", "cell_volume", "face_area"):
            assert np.array_equal(getattr(p, a), getattr(g, a)), a


@pytest.mark.gpu
def test_default_checkpoint_builds():
    if not os.path.exists(learned.CHECKPOINT):
        pytest.skip("no checkpoint committed")
    x = W.uniform(30000, 99)
    p = Voronoi.build(x, ZERO, ONE, 3, True)
    g = Voronoi.build(x, ZERO, ONE, 3, True, options=BuildOptions(mode=MODE_LEARNED))
    assert np.array_equal(p.face_right, g.face_right)
    assert np.array_equal(p.cell_volume, g.cell_volume)
